"""The reference's own RHLLC test on its own mesh, on the GPU.

tests/src/off_rhllc_test.cpp:22-45 of the reference: tests/data/brick.vtk
(118,783 tets) preprocessed by BuildDataFromVTK, GEOMETRY_TYPE Cone boundary
codes, initial state and inflow, CFL 0.7, h_min 1e-5, run to T = 0.5
(302 steps).  tests/golden/brick_cone.json holds the hashes of the compiled
reference's state at T = 0.5 (tests/golden/make_golden.py).  The FP64 GPU
step must reproduce it bit for bit, through the C ABI and through the
drop-in executables (rhllc_make_geo, then rhllc_solver); the drop-in's default
EOS TAUB run matches tests/golden/brick_cone_taub.json, the same test from the
reference built with its shipped EOS.
"""
import glob
import hashlib
import json
import lzma
import os
import re
import subprocess

import numpy as np
import pytest

from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M
from radrhd_solver_amd.build import MAKE_GEO_EXE, SOLVER_EXE

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def gold():
    with open(os.path.join(GOLD, "brick_cone.json")) as f:
        return json.load(f)


@pytest.fixture(scope="module")
def brick_vtk(tmp_path_factory):
    p = tmp_path_factory.mktemp("brick") / "brick.vtk"
    with lzma.open(os.path.join(GOLD, "brick.vtk.xz")) as f:
        p.write_bytes(f.read())
    return str(p)


@pytest.fixture(scope="module")
def brick_mesh(brick_vtk):
    pts, tets = M.read_vtk(brick_vtk)
    return M.bin_to_geo(M.tet_geometry(pts, tets, M.BOUND_CONE))


def _run(mesh, gold, precision=capi.PREC_FP64, guess=capi.GUESS_REFERENCE):
    W0 = np.tile(np.array(gold["W_init"]), (mesh.n_cells, 1))
    s = capi.Solver(mesh, cfl=gold["cfl"], h_min=gold["h_min"], precision=precision, guess=guess,
                    out_source=(gold["W_source"], gold["U_source"]))
    s.upload_W(W0)
    s.run(gold["steps"], graph=True)
    s.sync()
    W, U = s.download()
    stats = s.stats()
    s.close()
    return W, U, stats


def test_brick_reference_test_fp64_bitwise(brick_mesh, gold):
    assert sha(brick_mesh.cells) == gold["geo_cells_sha256"] and sha(brick_mesh.faces) == gold["geo_faces_sha256"]
    W, U, stats = _run(brick_mesh, gold)
    assert stats["steps"] == gold["steps"]
    assert stats["t"] == gold["t"] and stats["max_speed"] == gold["last_speed"]
    assert sha(U) == gold["U_sha256"]
    assert sha(W) == gold["W_sha256"]


def test_brick_reference_test_mixed_l1(brick_mesh, gold):
    """The mixed path on the same run: per-field relative L1 against FP64 <= 1e-5."""
    W64, U64, _ = _run(brick_mesh, gold)
    W, U, stats = _run(brick_mesh, gold, capi.PREC_MIXED, capi.GUESS_PREVIOUS)
    assert stats["fix_fail"] == 0
    for k in range(5):
        ref = np.abs(U64[:, k]).sum()
        assert np.abs(U[:, k] - U64[:, k]).sum() <= 1e-5 * ref, k


def _read(path, width):
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.int32)[0])
        return np.frombuffer(f.read(8 * n * width), np.float64).reshape(n, width)


@pytest.mark.parametrize("eos", ["ideal", None])
def test_brick_drop_in_executables(brick_vtk, gold, tmp_path, eos):
    """rhllc_make_geo <base> brick.vtk, then rhllc_solver with the test's
    configuration and no init file (the Cone default state).  Without an "eos"
    key the drop-in runs the reference's shipped EOS TAUB, checked against the
    reference built that way (tests/golden/brick_cone_taub.json)."""
    if eos is None:
        with open(os.path.join(GOLD, "brick_cone_taub.json")) as f:
            expect = json.load(f)
    else:
        expect = gold
    base = str(tmp_path) + os.sep
    r = subprocess.run([MAKE_GEO_EXE, base, brick_vtk], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    cfg = {"T": gold["T"], "CFL": gold["cfl"], "h_min": gold["h_min"], "save_timer": 0.1, "tau": 0.0}
    if eos is not None:
        cfg["eos"] = eos
    with open(base + "hllc_cfg.json", "w") as f:
        json.dump(cfg, f)
    with open(base + "solve_cfg.json", "w") as f:
        json.dump({"class_vtk": 0, "max_number_of_iter": 1, "accuracy": 1e-5, "start_point": 0}, f)
    os.makedirs(base + "Solve")
    keys = ["name_file_vtk", "name_file_sphere_direction", "base_address", "illum_geo_address", "graph_address",
            "trace_address", "solve_address", "add_dir_address", "tab_func_address", "name_file_hllc_set",
            "hllc_init_value", "solve_configuration"]
    dirs = {k: base for k in keys}
    dirs.update(name_file_vtk=brick_vtk, base_address=base, solve_address=base + "Solve/Solve",
                name_file_hllc_set=base + "hllc_cfg.json", hllc_init_value=base + "no_init_file.bin",
                solve_configuration=base + "solve_cfg.json")
    with open(base + "directories_cfg.json", "w") as f:
        json.dump(dirs, f)
    r = subprocess.run([SOLVER_EXE, base + "directories_cfg.json"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    ks = sorted(int(re.match(r".*Solve(\d+)density\.bin$", p).group(1)) for p in glob.glob(base + "Solve/Solve*density.bin"))
    assert ks[0] == 0 and len(ks) >= 3
    last = base + f"Solve/Solve{ks[-1]}"
    W = np.column_stack([_read(last + "density.bin", 1), _read(last + "velocity.bin", 3), _read(last + "pressure.bin", 1)])
    assert sha(W) == expect["W_sha256"]
    rho0 = _read(base + "Solve/Solve0density.bin", 1)
    assert (rho0 == gold["W_init"][0]).all()
