"""bench.py's host side, no GPU: the refusal of more than one GPU, the byte
figures its roofline uses, and the guard that takes the HBM traffic from
rocprofv3 counters only when they were measured on the same kernel sources
and mesh size."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402

N, F = 24118074, 48387834  # 159^3 hexes x 6 tets and their faces


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one MI355X" in r.stderr and not r.stdout


def test_step_model_bytes_per_cell_and_face():
    # the two-kernel step's model: FP64 face 80 N + 80 F, cell 153 N + 40 F; mixed face 20 N + 44 F,
    # cell 141 N + 20 F; the FP32 radiation fields 24 N in the cell kernel
    assert bench.model_bytes(1, 0, False, False) == (80, 153) and bench.model_bytes(0, 1, False, False) == (80, 40)
    assert bench.model_bytes(1, 0, True, False) == (20, 141) and bench.model_bytes(0, 1, True, False) == (44, 20)
    assert bench.model_bytes(1, 0, True, True) == (20, 165) and bench.model_bytes(1, 0, False, True) == (80, 177)
    assert sum(bench.model_bytes(N, F, True, True)) == 7_558_665_066
    assert sum(bench.model_bytes(N, F, False, True)) == 12_004_885_098


def test_tiled_design_bytes():
    # mixed tile: records 32 + U read/write 80 + 1/V 8 + slots 8 + records written 32 a cell,
    # 16 (S n) + 4 (far cell) a face; FP64 tile 176 a cell and 40 a face; radiation 24 a cell
    assert bench.design_bytes(1, 0, True, False, True) == 160 and bench.design_bytes(0, 1, True, False, True) == 20
    assert bench.design_bytes(1, 0, False, False, True) == 176 and bench.design_bytes(0, 1, False, False, True) == 40
    assert bench.design_bytes(N, F, True, True, True) == 5_405_482_296
    # every design figure stays below the model's, which still charges the flux array
    for mixed in (False, True):
        assert bench.design_bytes(N, F, mixed, True, True) < sum(bench.model_bytes(N, F, mixed, True))


def test_traffic_guard(monkeypatch):
    kernel = "k_tile<3, true, false, 0>"
    t, why = bench.load_traffic(kernel, N)
    assert t is not None, f"the committed rocprofv3 traffic no longer matches the kernel sources: {why}"
    assert 4e9 < t < 8e9
    assert bench.load_traffic("k_tile<2, true, false, 1>", N)[0] is not None  # the EOS TAUB line's kernel
    t, why = bench.load_traffic(kernel, N // 2)
    assert t is None and "cells" in why
    t, why = bench.load_traffic("k_unknown", N)
    assert t is None and "k_unknown" in why
    monkeypatch.setattr(bench, "kernel_source_sha", lambda: "0" * 16)
    t, why = bench.load_traffic(kernel, N)
    assert t is None and "other kernel sources" in why
