"""Loaders for the in-tree native libraries.

``libradrhd_mesh.so`` holds the host mesh/format tools; ``librhllc_hip.so``
holds the gfx950 HIP kernels behind the C ABI of ``include/rhllc_hip.h``.
Both are built in-tree by ``radrhd_solver_amd/build.py``; a missing library is
an error, never a reason to fall back to another code path.
"""
from __future__ import annotations

import ctypes as C
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
MESH_LIB = os.path.join(PKG_DIR, "libradrhd_mesh.so")
HIP_LIB = os.path.join(PKG_DIR, "librhllc_hip.so")

_mesh = None
_hip = None


def _need(path: str) -> None:
    if not os.path.exists(path):
        raise RuntimeError(
            f"native library {path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'`"
        )


def load_mesh_lib() -> C.CDLL:
    global _mesh
    if _mesh is None:
        _need(MESH_LIB)
        lib = C.CDLL(MESH_LIB)
        i64, vp, dbl, i32, u64 = C.c_int64, C.c_void_p, C.c_double, C.c_int32, C.c_uint64
        lib.rhllc_box_sizes.argtypes = [vp, vp, vp]
        lib.rhllc_box_build.argtypes = [vp, vp, vp]
        lib.rhllc_mesh_h_min.argtypes = [i64, vp, i64, vp]
        lib.rhllc_mesh_h_min.restype = dbl
        lib.rhllc_init_state.argtypes = [i32, i64, vp, vp, u64, vp]
        lib.rhllc_init_radiation.argtypes = [i64, u64, dbl, dbl, dbl, dbl, vp, vp, vp, vp]
        lib.rhllc_curve_order.argtypes = [i64, vp, i64, vp, vp, vp]
        lib.rhllc_file_count.argtypes = [C.c_char_p]
        lib.rhllc_file_count.restype = i64
        lib.rhllc_file_read_records.argtypes = [C.c_char_p, C.c_size_t, i64, vp]
        lib.rhllc_file_write_records.argtypes = [C.c_char_p, C.c_size_t, i64, vp]
        lib.rhllc_vtk_open.argtypes = [C.c_char_p, vp, vp, vp]
        lib.rhllc_vtk_copy.argtypes = [vp, vp, vp]
        lib.rhllc_vtk_close.argtypes = [vp]
        lib.rhllc_vtk_close.restype = None
        lib.rhllc_tet_geometry.argtypes = [i64, vp, i64, vp, i32, vp, vp, vp, vp, vp]
        lib.rhllc_bin_to_geo.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp, vp]
        lib.rhllc_table_read.argtypes = [C.c_char_p, vp, vp, vp, vp]
        lib.rhllc_mesh_last_error.restype = C.c_char_p
        _mesh = lib
    return _mesh


def load_hip_lib() -> C.CDLL:
    """The HIP solver library.  Loading needs no GPU; calls into it do."""
    global _hip
    if _hip is None:
        _need(HIP_LIB)
        from . import capi

        _hip = capi.declare(C.CDLL(HIP_LIB))
    return _hip
