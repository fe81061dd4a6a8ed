"""Synthetic meshes and the reference's binary geometry/state formats.

Record layouts are the reference's (lib/grids/include/solvers_struct.h):
``geo_face_t`` 40 B (:87-102), ``geo_cell_t`` 56 B (:120-154) and ``flux_t``
40 B (:51-70).  Files are "int32 n, then n records"
(lib/files_sys/include/reader_bin.h:113-135, :232-252).
"""
from __future__ import annotations

import ctypes as C
import json
import os
from dataclasses import dataclass

import numpy as np

from ._lib import load_mesh_lib

GEO_FACE_DTYPE = np.dtype(
    {
        "names": ["w0", "w1", "n", "S"],
        "formats": ["<u4", "<u4", ("<f8", 3), "<f8"],
        "offsets": [0, 4, 8, 32],
        "itemsize": 40,
    }
)
GEO_CELL_DTYPE = np.dtype(
    {
        "names": ["id_faces", "V", "center", "sign_n", "bits"],
        "formats": [("<i4", 4), "<f8", ("<f8", 3), "u1", "u1"],
        "offsets": [0, 16, 24, 48, 49],
        "itemsize": 56,
    }
)
FLUX_DTYPE = np.dtype((np.float64, 5))  # d, v[3], p

PERIODIC_X, PERIODIC_Y, PERIODIC_Z = 1, 2, 4
INIT_SHOCK_TUBE, INIT_BLAST, INIT_SHEAR, INIT_UNIFORM = 0, 1, 2, 3


class _BoxDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int32),
        ("ny", C.c_int32),
        ("nz", C.c_int32),
        ("origin", C.c_double * 3),
        ("length", C.c_double * 3),
        ("periodic", C.c_int32),
        ("jitter", C.c_double),
        ("seed", C.c_uint64),
        ("permute", C.c_int32),
    ]


def face_id_l(faces: np.ndarray) -> np.ndarray:
    return (faces["w0"] & 0x1FFFFFF).astype(np.int64)


def face_id_r(faces: np.ndarray) -> np.ndarray:
    """Sign-extended 26-bit id_r (negative values are boundary codes)."""
    v = (faces["w1"] & 0x3FFFFFF).astype(np.int64)
    return np.where(v >= (1 << 25), v - (1 << 26), v)


def cell_is_bound(cells: np.ndarray) -> np.ndarray:
    return (cells["bits"] >> 6) & 3


@dataclass
class Mesh:
    """Geometry records of an unstructured tetrahedral mesh."""

    cells: np.ndarray  # GEO_CELL_DTYPE[N]
    faces: np.ndarray  # GEO_FACE_DTYPE[F]

    @property
    def n_cells(self) -> int:
        return int(self.cells.shape[0])

    @property
    def n_faces(self) -> int:
        return int(self.faces.shape[0])

    def h_min(self) -> float:
        lib = load_mesh_lib()
        h = float(lib.rhllc_mesh_h_min(self.n_cells, self.cells.ctypes.data, self.n_faces, self.faces.ctypes.data))
        if h < 0:
            raise ValueError(lib.rhllc_mesh_last_error().decode())
        return h

    def write(self, faces_path: str, cells_path: str) -> None:
        write_records(faces_path, self.faces)
        write_records(cells_path, self.cells)

    @staticmethod
    def read(faces_path: str, cells_path: str) -> "Mesh":
        return Mesh(read_records(cells_path, GEO_CELL_DTYPE), read_records(faces_path, GEO_FACE_DTYPE))


def kuhn_box(
    nx: int,
    ny: int,
    nz: int,
    origin=(0.0, 0.0, 0.0),
    length=(1.0, 1.0, 1.0),
    periodic: int = 0,
    jitter: float = 0.05,
    seed: int = 20261202,
    permute: bool = True,
) -> Mesh:
    """A box of nx*ny*nz hexes cut into 6 Kuhn tets each, jittered and permuted."""
    lib = load_mesh_lib()
    d = _BoxDesc(nx, ny, nz, (C.c_double * 3)(*origin), (C.c_double * 3)(*length), periodic, jitter, seed, int(permute))
    n_c, n_f = C.c_int64(), C.c_int64()
    if lib.rhllc_box_sizes(C.byref(d), C.byref(n_c), C.byref(n_f)) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    cells = np.zeros(n_c.value, dtype=GEO_CELL_DTYPE)
    faces = np.zeros(n_f.value, dtype=GEO_FACE_DTYPE)
    if lib.rhllc_box_build(C.byref(d), cells.ctypes.data, faces.ctypes.data) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return Mesh(cells, faces)


def init_state(mesh: Mesh, kind: int, params=None, seed: int = 20261202) -> np.ndarray:
    """Initial primitive state W[N,5] = (rho, vx, vy, vz, p)."""
    lib = load_mesh_lib()
    W = np.zeros((mesh.n_cells, 5), dtype=np.float64)
    p = None
    if params is not None:
        arr = np.ascontiguousarray(params, dtype=np.float64)
        p = arr.ctypes.data
    if lib.rhllc_init_state(kind, mesh.n_cells, mesh.cells.ctypes.data, p, seed, W.ctypes.data) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return W


def init_radiation(n_cells: int, seed: int = 20261202, ur=(1e-3, 1.0), kappa=(0.1, 10.0)):
    """Seeded FP32 radiation fields (Ur, Fr[N,3], kappa_a, kappa_s)."""
    lib = load_mesh_lib()
    Ur = np.zeros(n_cells, np.float32)
    Fr = np.zeros((n_cells, 3), np.float32)
    ka = np.zeros(n_cells, np.float32)
    ks = np.zeros(n_cells, np.float32)
    rc = lib.rhllc_init_radiation(
        n_cells, seed, ur[0], ur[1], kappa[0], kappa[1], Ur.ctypes.data, Fr.ctypes.data, ka.ctypes.data, ks.ctypes.data
    )
    if rc != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return Ur, Fr, ka, ks


def curve_order(mesh: Mesh):
    """(cell_new_to_old, face_new_to_old) relabelling permutations (Hilbert order)."""
    lib = load_mesh_lib()
    cp = np.zeros(mesh.n_cells, np.int32)
    fp = np.zeros(mesh.n_faces, np.int32)
    rc = lib.rhllc_curve_order(
        mesh.n_cells, mesh.cells.ctypes.data, mesh.n_faces, mesh.faces.ctypes.data, cp.ctypes.data, fp.ctypes.data
    )
    if rc != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return cp, fp


BOUND_CONE, BOUND_FREE = 0, 1


def read_vtk(path: str):
    """(points[P,3] float64, tets[N,4] int32) of a legacy-format VTK tet grid."""
    lib = load_mesh_lib()
    h = C.c_void_p()
    n_p, n_c = C.c_int64(), C.c_int64()
    if lib.rhllc_vtk_open(os.fsencode(path), C.byref(h), C.byref(n_p), C.byref(n_c)) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    try:
        points = np.zeros((n_p.value, 3), np.float64)
        tets = np.zeros((n_c.value, 4), np.int32)
        if lib.rhllc_vtk_copy(h, points.ctypes.data, tets.ctypes.data) != 0:
            raise ValueError(lib.rhllc_mesh_last_error().decode())
    finally:
        lib.rhllc_vtk_close(h)
    return points, tets


def read_table(path: str):
    """(header[6], data[nx, ny]) of a cooling-function table (tab_func/rcf_H.dat format);
    header = (min_x, max_x, step_x, min_y, max_y, step_y) over (ln D, ln T)."""
    lib = load_mesh_lib()
    nx, ny = C.c_int32(), C.c_int32()
    hdr = np.zeros(6)
    if lib.rhllc_table_read(os.fsencode(path), C.byref(nx), C.byref(ny), hdr.ctypes.data, None) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    data = np.zeros((nx.value, ny.value))
    if lib.rhllc_table_read(os.fsencode(path), C.byref(nx), C.byref(ny), hdr.ctypes.data, data.ctypes.data) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return hdr, data


@dataclass
class TetGeometry:
    """The per-cell arrays of the reference's preprocessing (BuildDataFromVTK,
    lib/geometry/src/make_internal_format.cpp:272-336)."""

    neighbors: np.ndarray  # int32[4N]: neighbour cell*4 + local face, or a boundary code
    normals: np.ndarray  # float64[N,4,3] outward unit normals
    areas: np.ndarray  # float64[4N]
    volumes: np.ndarray  # float64[N]
    centers: np.ndarray  # float64[N,3]

    def write(self, base: str) -> None:
        """neighbors.bin, normals.bin, areas.bin, volume.bin, centers.bin
        (lib/global_value.h:74-78), each "int32 n, then n records"."""
        write_records(base + "neighbors.bin", self.neighbors)
        write_records(base + "normals.bin", self.normals)
        write_records(base + "areas.bin", self.areas)
        write_records(base + "volume.bin", self.volumes)
        write_records(base + "centers.bin", self.centers)

    @staticmethod
    def read(base: str) -> "TetGeometry":
        return TetGeometry(
            read_records(base + "neighbors.bin", np.int32),
            read_records(base + "normals.bin", np.dtype((np.float64, (4, 3)))),
            read_records(base + "areas.bin", np.float64),
            read_records(base + "volume.bin", np.float64),
            read_records(base + "centers.bin", np.dtype((np.float64, 3))),
        )


def tet_geometry(points: np.ndarray, tets: np.ndarray, bound: int = BOUND_CONE) -> TetGeometry:
    lib = load_mesh_lib()
    points = np.ascontiguousarray(points, np.float64)
    tets = np.ascontiguousarray(tets, np.int32)
    n = tets.shape[0]
    g = TetGeometry(
        np.zeros(4 * n, np.int32),
        np.zeros((n, 4, 3), np.float64),
        np.zeros(4 * n, np.float64),
        np.zeros(n, np.float64),
        np.zeros((n, 3), np.float64),
    )
    rc = lib.rhllc_tet_geometry(
        points.shape[0], points.ctypes.data, n, tets.ctypes.data, bound, g.neighbors.ctypes.data,
        g.normals.ctypes.data, g.areas.ctypes.data, g.volumes.ctypes.data, g.centers.ctypes.data,
    )
    if rc != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return g


def bin_to_geo(g: TetGeometry) -> Mesh:
    """ConvertBinToGeo (lib/geometry/src/convert_bin_to_geo.cpp:3-48)."""
    lib = load_mesh_lib()
    n = g.volumes.shape[0]
    arrs = [np.ascontiguousarray(a) for a in (g.neighbors, g.normals, g.areas, g.volumes, g.centers)]
    ptrs = [a.ctypes.data for a in arrs]
    n_f = C.c_int64()
    if lib.rhllc_bin_to_geo(n, *ptrs, C.byref(n_f), None, None) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    cells = np.zeros(n, GEO_CELL_DTYPE)
    faces = np.zeros(n_f.value, GEO_FACE_DTYPE)
    if lib.rhllc_bin_to_geo(n, *ptrs, C.byref(n_f), cells.ctypes.data, faces.ctypes.data) != 0:
        raise ValueError(lib.rhllc_mesh_last_error().decode())
    return Mesh(cells, faces)


def build_geo_from_vtk(vtk_path: str, base: str, bound: int = BOUND_CONE):
    """make_all_geo_data's RHLLC part (utils/src/make_all_geo_data.cpp:12-41):
    BuildDataFromVTK's five arrays, then BinToGeo's geo_faces.bin and
    geo_cells.bin, all written under the directory prefix `base`."""
    points, tets = read_vtk(vtk_path)
    g = tet_geometry(points, tets, bound)
    g.write(base)
    m = bin_to_geo(g)
    m.write(base + "geo_faces.bin", base + "geo_cells.bin")
    return g, m


def write_records(path: str, recs: np.ndarray) -> None:
    with open(path, "wb") as f:
        f.write(np.int32(recs.shape[0]).tobytes())
        f.write(np.ascontiguousarray(recs).tobytes())


def read_records(path: str, dtype) -> np.ndarray:
    dtype = np.dtype(dtype)
    with open(path, "rb") as f:
        n = int(np.frombuffer(f.read(4), np.int32)[0])
        data = f.read(n * dtype.itemsize)
    if len(data) != n * dtype.itemsize:
        raise IOError(f"short read in {path}")
    # A writable view of the raw bytes (a structured copy would not keep the
    # records' padding bytes).
    return np.frombuffer(bytearray(data), dtype=dtype)


def write_hllc_cfg(path: str, T: float, CFL: float, h_min: float, save_timer: float, tau: float = 0.0,
                   eos: str | None = None, geometry: str | None = None, extra: dict | None = None) -> None:
    """hllc_cfg.json with the keys of lib/json/include/json_struct.h:15-20, and
    the drop-in's optional "eos" ("taub" when absent, the reference's shipped EOS),
    "geometry" (the reference's GEOMETRY_TYPE, "cone" when absent) and any other
    of its keys in `extra` ("precision", "c2p_start", "mlp_model", "gamma")."""
    cfg = {"T": T, "CFL": CFL, "h_min": h_min, "save_timer": save_timer, "tau": tau}
    if eos is not None:
        cfg["eos"] = eos
    if geometry is not None:
        cfg["geometry"] = geometry
    cfg.update(extra or {})
    with open(path, "w") as f:
        json.dump(cfg, f, indent=4)


def write_case(dirpath: str, mesh: Mesh, W: np.ndarray, T: float, CFL: float, save_timer: float,
               eos: str | None = None, geometry: str | None = None, extra: dict | None = None) -> dict:
    """Write geo_faces.bin, geo_cells.bin, the flux_t initial state and the three JSON configs."""
    os.makedirs(dirpath, exist_ok=True)
    base = os.path.abspath(dirpath) + os.sep
    mesh.write(base + "geo_faces.bin", base + "geo_cells.bin")
    write_records(base + "hllc_init.bin", np.ascontiguousarray(W, dtype=np.float64))
    write_hllc_cfg(base + "hllc_cfg.json", T, CFL, mesh.h_min(), save_timer, eos=eos, geometry=geometry, extra=extra)
    with open(base + "solve_cfg.json", "w") as f:
        json.dump({"class_vtk": 0, "max_number_of_iter": 1, "accuracy": 1e-5, "start_point": 0}, f, indent=4)
    os.makedirs(base + "Solve", exist_ok=True)
    dirs = {
        "name_file_vtk": base + "mesh.vtk",
        "name_file_sphere_direction": base + "grid_direction.txt",
        "base_address": base,
        "illum_geo_address": base + "illum_geo/",
        "graph_address": base + "graph/",
        "trace_address": base + "trace/",
        "solve_address": base + "Solve/Solve",
        "add_dir_address": base + "add_dir/",
        "tab_func_address": base + "tab_func/",
        "name_file_hllc_set": base + "hllc_cfg.json",
        "hllc_init_value": base + "hllc_init.bin",
        "solve_configuration": base + "solve_cfg.json",
    }
    with open(base + "directories_cfg.json", "w") as f:
        json.dump(dirs, f, indent=4)
    return dirs
