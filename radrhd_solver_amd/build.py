"""In-tree build of the native libraries (no JIT cache, no pip install).

* ``libradrhd_mesh.so``  -- host mesh/format tools (g++, OpenMP).
* ``librhllc_hip.so``    -- gfx950 HIP kernels + C ABI (hipcc).

HIP code is compiled with ``-ffp-contract=off`` (the FP64 path is bitwise
comparable with the reference) and with packed-FP32 instructions disabled.
"""
from __future__ import annotations

import os
import subprocess
import sys

from ._lib import HIP_LIB, MESH_LIB, PKG_DIR

CSRC = os.path.join(PKG_DIR, "csrc")
MESH_SRC = [
    os.path.join(CSRC, "mesh", f)
    for f in ("kuhn_box.cpp", "mesh_tools.cpp", "curve_order.cpp", "mesh_check.cpp", "vtk_legacy.cpp",
              "tet_geometry.cpp", "table_txt.cpp")
]
HIP_SRC = [os.path.join(CSRC, "hip", f) for f in ("rhllc_kernels.hip", "rhllc_capi.hip", "mlp_kernels.hip", "mlp_capi.hip")]
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def _run(cmd):
    print("+", " ".join(cmd), flush=True)
    subprocess.run(cmd, check=True)


def build_mesh() -> str:
    # -ffp-contract=off: the mesh geometry is compared bitwise with the reference's.
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared", "-Wall", "-o", MESH_LIB]
         + MESH_SRC)
    return MESH_LIB


def build_hip() -> str:
    objdir = os.path.join(PKG_DIR, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    for src in HIP_SRC:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        _run([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj])
        objs.append(obj)
    # Host preparation of a handle's layout (validation, relabelling, tiles):
    # g++ with OpenMP (libgomp, as the mesh library uses).
    for src in (os.path.join(CSRC, "mesh", "curve_order.cpp"), os.path.join(CSRC, "mesh", "mesh_check.cpp"),
                os.path.join(CSRC, "hip", "rhllc_layout.cpp")):
        obj = os.path.join(objdir, os.path.splitext(os.path.basename(src))[0] + ".o")
        _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-c", src, "-o", obj])
        objs.append(obj)
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


SHIM_DIR = os.path.join(PKG_DIR, "shim")
SOLVER_EXE = os.path.join(PKG_DIR, "bin", "rhllc_solver")
MAKE_GEO_EXE = os.path.join(PKG_DIR, "bin", "rhllc_make_geo")
SOLVER_MPI_EXE = os.path.join(PKG_DIR, "bin", "rhllc_solver_mpi")
DROPIN_LIB = os.path.join(PKG_DIR, "librhllc_dropin.so")


def build_shim() -> str:
    """The drop-in C++ interface (librhllc_dropin.so) and the rhllc_solver,
    rhllc_solver_mpi (one rank) and rhllc_make_geo executables."""
    os.makedirs(os.path.dirname(SOLVER_EXE), exist_ok=True)
    inc = ["-I", os.path.join(os.path.dirname(PKG_DIR), "include")]
    link = ["-L", PKG_DIR, "-lrhllc_hip", "-lradrhd_mesh", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,$ORIGIN/.."]
    _run(["g++", "-std=c++17", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-Wall", "-Wno-invalid-offsetof",
          "-o", DROPIN_LIB, os.path.join(SHIM_DIR, "rhllc_dropin.cpp")] + inc + link)
    exe_link = ["-L", PKG_DIR, "-lrhllc_dropin", "-lrhllc_hip", "-lradrhd_mesh", "-Wl,-rpath,$ORIGIN/.."]
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-o", SOLVER_EXE, os.path.join(SHIM_DIR, "rhllc_solver_main.cpp")]
         + inc + exe_link)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-o", SOLVER_MPI_EXE, os.path.join(SHIM_DIR, "rhllc_solver_mpi_main.cpp")]
         + inc + exe_link)
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-o", MAKE_GEO_EXE, os.path.join(SHIM_DIR, "rhllc_make_geo_main.cpp")]
         + inc + exe_link)
    return SOLVER_EXE


def build_all() -> None:
    build_mesh()
    build_hip()
    build_shim()


if __name__ == "__main__":
    build_all()
    sys.exit(0)
