"""Test oracle for the RHLLC step.

TEST INFRASTRUCTURE ONLY.  Only tests/, ``__graft_entry__.smoke()`` and the
``cpu_baseline`` leg of bench.py may import this package, and only as the
checker.  It wraps:

* ``liboracle_rhllc.so`` -- the C restatement in ``rhllc_oracle.c`` of the
  reference's ``rhllc::Hllc3dStab`` and its callees (FP64, AoS records);
* ``_ref/libref_rhllc.so`` and ``_ref/libref_rhllc_taub.so`` -- the
  reference's own sources compiled by ``make ref`` with EOS IDEAL and with
  EOS TAUB (present only where the reference tree was available at build
  time).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from dataclasses import dataclass

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(HERE, "liboracle_rhllc.so")
ORACLE_FAST_LIB = os.path.join(HERE, "liboracle_rhllc_fast.so")  # -O3 timing build (cpu_baseline only)
FAST_FLAGS = "gcc -std=c11 -O3 -fopenmp -fPIC"
REF_LIB = os.path.join(HERE, "_ref", "libref_rhllc.so")
REF_LIB_TAUB = os.path.join(HERE, "_ref", "libref_rhllc_taub.so")

FLUX = np.dtype((np.float64, 5))
ELEM_BYTES, FACE_BYTES = 136, 80


class Physics(C.Structure):
    _fields_ = [
        ("gamma", C.c_double),
        ("has_out_source", C.c_int),
        ("out_source_W", C.c_double * 5),
        ("out_source_U", C.c_double * 5),
        ("eos", C.c_int),
        ("lock_wall", C.c_int),
        ("has_inner_source", C.c_int),
        ("inner_source_W", C.c_double * 5),
        ("inner_source_U", C.c_double * 5),
    ]


EOS_IDEAL, EOS_TAUB = 0, 1  # ORC_EOS_* of rhllc_oracle.h


class Stats(C.Structure):
    _fields_ = [
        ("c2p_fail", C.c_int64 * 5),
        ("fix_fail", C.c_int64),
        ("lor_die", C.c_int64),
        ("newton_iters", C.c_int64),
    ]

    def as_dict(self):
        return {
            "c2p_fail": list(self.c2p_fail)[1:],
            "fix_fail": self.fix_fail,
            "lor_die": self.lor_die,
            "newton_iters": self.newton_iters,
        }


_orc = None
_ref = {}


def build(ref: bool = True) -> None:
    """Compile the oracle (and the reference build when the tree is present)."""
    subprocess.run(["make", "-s", "-C", HERE, "liboracle_rhllc.so", "liboracle_rhllc_fast.so"], check=True)
    if ref and os.path.isdir("/root/reference/solvers/rhllc"):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _p(a):
    return None if a is None else a.ctypes.data


_orc_fast = None


def lib(fast: bool = False):
    """The strict build (parity), or with fast=True the -O3 timing build."""
    global _orc, _orc_fast
    if fast:
        if _orc_fast is None:
            _orc_fast = _declare(ORACLE_FAST_LIB)
        return _orc_fast
    if _orc is None:
        _orc = _declare(ORACLE_LIB)
    return _orc


def _declare(path):
    if True:
        if not os.path.exists(path):
            raise RuntimeError(f"{path} is not built (make -C oracle)")
        L = C.CDLL(path)
        vp, i64, dbl, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
        L.orc_p2c_n.argtypes = [vp, i64, vp, vp]
        L.orc_c2p_n.argtypes = [vp, i64, vp, vp, vp, vp]
        L.orc_pressure_fix_n.argtypes = [vp, i64, vp, vp, vp]
        L.orc_rotation.argtypes = [vp, vp]
        L.orc_wall_mirror.argtypes = [vp, vp]
        L.orc_flux_n.argtypes = [vp, i64, vp, vp, vp, vp, vp, vp, vp]
        L.orc_load.argtypes = [i64, vp, vp, i64, vp, vp, vp]
        L.orc_step.argtypes = [vp, i64, vp, i64, vp, dbl, vp]
        L.orc_step.restype = dbl
        L.orc_run.argtypes = [vp, i64, vp, i64, vp, dbl, dbl, i32, vp, vp, vp]
        L.orc_run.restype = dbl
        L.orc_edd_tensor.argtypes = [i32, i32, dbl, vp]
        L.orc_edd_tensor.restype = dbl
        L.orc_rad_source.argtypes = [vp, dbl, vp, dbl, dbl, i32, vp, vp]
        L.orc_run_rad.argtypes = [vp, i64, vp, i64, vp, dbl, dbl, i32, vp, vp, vp, vp, i32, vp, vp, vp, vp, vp]
        L.orc_run_rad.restype = dbl
        L.orc_step_rad.argtypes = [vp, i64, vp, i64, vp, dbl, vp, vp, vp, vp, i32, vp, vp, vp]
        L.orc_step_rad.restype = dbl
        L.orc_table_read.argtypes = [C.c_char_p, vp, vp, vp, vp]
        L.orc_table_read.restype = i32
        L.orc_table_lookup.argtypes = [vp, dbl, dbl]
        L.orc_table_lookup.restype = dbl
        L.orc_cell_rad.argtypes = [vp, vp, vp]
        L.orc_rad_source_opt.argtypes = [vp, vp, dbl, vp, vp, vp]
        L.orc_num_threads.restype = i32
        return L


def _ref_path(eos):
    return REF_LIB_TAUB if eos == EOS_TAUB else REF_LIB


def have_ref(eos=EOS_IDEAL) -> bool:
    return os.path.exists(_ref_path(eos))


def ref_lib(eos=EOS_IDEAL):
    """The reference compiled with EOS IDEAL (default) or EOS TAUB."""
    if eos not in _ref:
        if not have_ref(eos):
            raise RuntimeError(f"{_ref_path(eos)} is not built (make -C oracle ref)")
        L = C.CDLL(_ref_path(eos))
        vp, i64, dbl, i32 = C.c_void_p, C.c_int64, C.c_double, C.c_int
        L.ref_layout.argtypes = [vp]
        L.ref_gamma.restype = dbl
        L.ref_set_log_dir.argtypes = [C.c_char_p]
        L.ref_p2c_n.argtypes = [i64, vp, vp]
        L.ref_c2p_n.argtypes = [i64, vp, vp, vp]
        L.ref_pressure_fix_n.argtypes = [i64, vp, vp, vp]
        L.ref_rotation_n.argtypes = [i64, vp, vp]
        L.ref_flux_n.argtypes = [i64, vp, vp, vp, vp, vp, vp, vp]
        L.ref_run.argtypes = [i64, vp, vp, i64, vp, dbl, dbl, i32, vp, vp, vp, vp]
        L.ref_convert_bin_to_geo.argtypes = [i64, vp, vp, vp, vp, vp, i64, vp, vp]
        L.ref_convert_bin_to_geo.restype = i64
        L.ref_cone_states.argtypes = [vp, vp, vp]
        L.ref_table_read.argtypes = [C.c_char_p, vp, vp, vp, vp]
        L.ref_table_read.restype = i32
        L.ref_table_lookup.argtypes = [i32, i32, vp, vp, i64, vp, vp, vp]
        L.ref_set_log_dir(("/tmp/").encode())
        _ref[eos] = L
    return _ref[eos]


def physics(gamma: float = 5.0 / 3.0, eos: int = EOS_IDEAL) -> Physics:
    ph = Physics()
    ph.gamma = gamma
    ph.eos = eos
    return ph


# ---------------------------------------------------------------- oracle calls
def p2c(W, gamma=5.0 / 3.0, eos=EOS_IDEAL):
    """GetConvValueStab over n states; returns (W_clamped, U)."""
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    U = np.zeros_like(W)
    ph = physics(gamma, eos)
    lib().orc_p2c_n(C.byref(ph), W.shape[0], W.ctypes.data, U.ctypes.data)
    return W, U


def c2p(U, gamma=5.0 / 3.0, W_init=None, eos=EOS_IDEAL):
    """GetPhysValueStab over n states; returns (W, codes, iters).  W keeps W_init where c2p fails."""
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 5)
    W = np.zeros_like(U) if W_init is None else np.array(W_init, dtype=np.float64, order="C").reshape(-1, 5)
    codes = np.zeros(U.shape[0], np.int32)
    iters = np.zeros(U.shape[0], np.int32)
    ph = physics(gamma, eos)
    lib().orc_c2p_n(C.byref(ph), U.shape[0], U.ctypes.data, W.ctypes.data, codes.ctypes.data, iters.ctypes.data)
    return W, codes, iters


def pressure_fix(U, W, gamma=5.0 / 3.0, eos=EOS_IDEAL):
    U = np.array(U, dtype=np.float64, order="C").reshape(-1, 5)
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    codes = np.zeros(U.shape[0], np.int32)
    ph = physics(gamma, eos)
    lib().orc_pressure_fix_n(C.byref(ph), U.shape[0], U.ctypes.data, W.ctypes.data, codes.ctypes.data)
    return U, W, codes


def rotation(n):
    n = np.ascontiguousarray(n, dtype=np.float64).reshape(3)
    T = np.zeros(9)
    lib().orc_rotation(n.ctypes.data, T.ctypes.data)
    return T.reshape(3, 3)


def wall_mirror(n, v):
    """The e_bound_lock reflection of v about the face normal n (rhllc_utils.cpp:95-110)."""
    n = np.ascontiguousarray(n, dtype=np.float64).reshape(3)
    v = np.array(v, dtype=np.float64).reshape(3)
    lib().orc_wall_mirror(n.ctypes.data, v.ctypes.data)
    return v


def flux(UL, UR, WL, WR, nS4, gamma=5.0 / 3.0, eos=EOS_IDEAL):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5) for a in (UL, UR, WL, WR)]
    nS4 = np.ascontiguousarray(nS4, dtype=np.float64).reshape(-1, 4)
    n = nS4.shape[0]
    F = np.zeros((n, 5))
    speed = np.zeros(n)
    ph = physics(gamma, eos)
    lib().orc_flux_n(C.byref(ph), n, *[a.ctypes.data for a in arrs], nS4.ctypes.data, F.ctypes.data, speed.ctypes.data)
    return F, speed


class State:
    """AoS elem_t / face_t records of the oracle for one mesh."""

    def __init__(self, mesh, W0, gamma=5.0 / 3.0, eos=EOS_IDEAL, fast=False):
        self.L = lib(fast)
        self.n_cells, self.n_faces = mesh.n_cells, mesh.n_faces
        self.cells = np.zeros(self.n_cells * ELEM_BYTES, np.uint8)
        self.faces = np.zeros(self.n_faces * FACE_BYTES, np.uint8)
        W0 = np.ascontiguousarray(W0, dtype=np.float64)
        self.L.orc_load(
            self.n_cells, mesh.cells.ctypes.data, W0.ctypes.data, self.n_faces, mesh.faces.ctypes.data,
            self.cells.ctypes.data, self.faces.ctypes.data,
        )
        self.ph = physics(gamma, eos)
        self.stats = Stats()

    def _view(self, off):
        v = self.cells.view(np.float64).reshape(self.n_cells, ELEM_BYTES // 8)
        return v[:, off // 8: off // 8 + 5]

    @property
    def W(self):
        return np.array(self._view(0))

    @property
    def U(self):
        return np.array(self._view(40))

    @property
    def face_flux(self):
        return np.array(self.faces.view(np.float64).reshape(self.n_faces, FACE_BYTES // 8)[:, :5])

    def run(self, cfl, h_min, steps):
        taus = np.zeros(steps)
        speeds = np.zeros(steps)
        t = self.L.orc_run(
            C.byref(self.ph), self.n_cells, self.cells.ctypes.data, self.n_faces, self.faces.ctypes.data,
            cfl, h_min, steps, taus.ctypes.data, speeds.ctypes.data, C.byref(self.stats),
        )
        return t, taus, speeds

    def run_rad(self, cfl, h_min, steps, Ur, Fr, ka=None, ks=None, emission=0, table=None, Tr=None):
        """Coupled steps; with `table` (a Table) the opacities come from each cell's state; with
        Tr (n x 3 x 3, FP32) the pressure-tensor branch instead of the M1 closure."""
        taus = np.zeros(steps)
        speeds = np.zeros(steps)
        if ka is None:
            ka = ks = np.zeros(self.n_cells, np.float32)
        Ur, Fr, ka, ks = (np.ascontiguousarray(a, dtype=np.float32) for a in (Ur, Fr, ka, ks))
        T = None if Tr is None else np.ascontiguousarray(Tr, dtype=np.float32).reshape(self.n_cells, 9)
        tab = None if table is None else C.byref(table.c())
        t = self.L.orc_run_rad(
            C.byref(self.ph), self.n_cells, self.cells.ctypes.data, self.n_faces, self.faces.ctypes.data,
            cfl, h_min, steps, Ur.ctypes.data, Fr.ctypes.data, ka.ctypes.data, ks.ctypes.data, emission,
            tab, _p(T), taus.ctypes.data, speeds.ctypes.data, C.byref(self.stats),
        )
        return t, taus, speeds

    def step_rad(self, tau, Ur, Fr, ka=None, ks=None, emission=0, table=None, Tr=None):
        """One coupled step (hydro, source, c2p); returns the step's max signal speed."""
        if ka is None:
            ka = ks = np.zeros(self.n_cells, np.float32)
        Ur, Fr, ka, ks = (np.ascontiguousarray(a, dtype=np.float32) for a in (Ur, Fr, ka, ks))
        T = None if Tr is None else np.ascontiguousarray(Tr, dtype=np.float32).reshape(self.n_cells, 9)
        tab = None if table is None else C.byref(table.c())
        return self.L.orc_step_rad(
            C.byref(self.ph), self.n_cells, self.cells.ctypes.data, self.n_faces, self.faces.ctypes.data, tau,
            Ur.ctypes.data, Fr.ctypes.data, ka.ctypes.data, ks.ctypes.data, emission, tab, _p(T),
            C.byref(self.stats),
        )

    def step(self, tau):
        return self.L.orc_step(
            C.byref(self.ph), self.n_cells, self.cells.ctypes.data, self.n_faces, self.faces.ctypes.data, tau,
            C.byref(self.stats),
        )


def rad_source(W, Ur, Fr, ka, ks, emission=0, Tr=None):
    W = np.ascontiguousarray(W, dtype=np.float64).reshape(5)
    Fr = np.ascontiguousarray(Fr, dtype=np.float64).reshape(3)
    T = None if Tr is None else np.ascontiguousarray(Tr, dtype=np.float32).reshape(9)
    G = np.zeros(4)
    lib().orc_rad_source(W.ctypes.data, float(Ur), Fr.ctypes.data, float(ka), float(ks), emission, _p(T),
                         G.ctypes.data)
    return G


class _CTable(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32)] + [
        (k, C.c_double) for k in ("min_x", "max_x", "step_x", "min_y", "max_y", "step_y")
    ] + [("data", C.c_void_p)]


@dataclass
class Table:
    """TableFunc (ln Lambda over ln D x ln T): hdr = (min_x, max_x, step_x, min_y, max_y, step_y)."""

    hdr: np.ndarray
    data: np.ndarray  # (nx, ny) float64

    def c(self):
        self.data = np.ascontiguousarray(self.data, dtype=np.float64)
        nx, ny = self.data.shape
        self._c = _CTable(nx, ny, *[float(v) for v in self.hdr], self.data.ctypes.data)
        return self._c


def table_read(path):
    """ReadTableFunc restated (orc_table_read)."""
    nx, ny = C.c_int32(), C.c_int32()
    hdr = np.zeros(6)
    if lib().orc_table_read(str(path).encode(), C.byref(nx), C.byref(ny), hdr.ctypes.data, None):
        raise ValueError(f"cannot read table {path}")
    data = np.zeros((nx.value, ny.value))
    lib().orc_table_read(str(path).encode(), C.byref(nx), C.byref(ny), hdr.ctypes.data, data.ctypes.data)
    return Table(hdr, data)


def table_lookup(t, x, y):
    ct = t.c()
    x, y = np.broadcast_arrays(np.asarray(x, np.float64), np.asarray(y, np.float64))
    return np.array([lib().orc_table_lookup(C.byref(ct), float(a), float(b)) for a, b in zip(x.ravel(), y.ravel())])


def cell_rad(t, W):
    """full_phys_data_t::Init per state: (T, vel, lorenz, alpha, betta), shape (n, 5)."""
    W = np.ascontiguousarray(np.atleast_2d(W), dtype=np.float64)
    ct = t.c()
    out = np.zeros((len(W), 5))
    for i in range(len(W)):
        lib().orc_cell_rad(C.byref(ct), W[i].ctypes.data, out[i].ctypes.data)
    return out


def rad_source_opt(W, cell, Ur, Fr, Tr=None):
    W = np.ascontiguousarray(W, dtype=np.float64).reshape(5)
    cell = np.ascontiguousarray(cell, dtype=np.float64).reshape(5)
    Fr = np.ascontiguousarray(Fr, dtype=np.float64).reshape(3)
    T = None if Tr is None else np.ascontiguousarray(Tr, dtype=np.float32).reshape(9)
    G = np.zeros(4)
    lib().orc_rad_source_opt(W.ctypes.data, cell.ctypes.data, float(Ur), Fr.ctypes.data, _p(T), G.ctypes.data)
    return G


# ------------------------------------------------------------- reference calls
def ref_run(mesh, W0, cfl, h_min, steps, eos=EOS_IDEAL):
    L = ref_lib(eos)
    W0 = np.ascontiguousarray(W0, dtype=np.float64)
    Wo = np.zeros((mesh.n_cells, 5))
    Uo = np.zeros((mesh.n_cells, 5))
    taus = np.zeros(steps)
    speeds = np.zeros(steps)
    L.ref_run(
        mesh.n_cells, mesh.cells.ctypes.data, W0.ctypes.data, mesh.n_faces, mesh.faces.ctypes.data, cfl, h_min,
        steps, Wo.ctypes.data, Uo.ctypes.data, taus.ctypes.data, speeds.ctypes.data,
    )
    return Wo, Uo, taus, speeds


def ref_convert_bin_to_geo(nb, normals, areas, vols, cents):
    """The reference's ConvertBinToGeo: (cells bytes [N*56], faces bytes [F*40])."""
    L = ref_lib()
    arrs = [np.ascontiguousarray(nb, np.int32)] + [np.ascontiguousarray(a, np.float64) for a in (normals, areas, vols, cents)]
    n = arrs[3].shape[0]
    cells = np.zeros(n * 56, np.uint8)
    nf = L.ref_convert_bin_to_geo(n, *[a.ctypes.data for a in arrs], 0, cells.ctypes.data, None)
    faces = np.zeros(nf * 40, np.uint8)
    L.ref_convert_bin_to_geo(n, *[a.ctypes.data for a in arrs], nf, cells.ctypes.data, faces.ctypes.data)
    return cells, faces


def ref_cone_states():
    """(W_init, W_source, U_source) of the reference build's GEOMETRY_TYPE Cone."""
    W0, Ws, Us = np.zeros(5), np.zeros(5), np.zeros(5)
    ref_lib().ref_cone_states(W0.ctypes.data, Ws.ctypes.data, Us.ctypes.data)
    return W0, Ws, Us


def ref_table_read(path):
    """files_sys::txt::ReadTableFunc of the compiled reference."""
    nx, ny = C.c_int32(), C.c_int32()
    hdr = np.zeros(6)
    if ref_lib().ref_table_read(str(path).encode(), C.byref(nx), C.byref(ny), hdr.ctypes.data, None):
        raise ValueError(f"cannot read table {path}")
    data = np.zeros((nx.value, ny.value))
    ref_lib().ref_table_read(str(path).encode(), C.byref(nx), C.byref(ny), hdr.ctypes.data, data.ctypes.data)
    return Table(hdr, data)


def ref_table_lookup(t, x, y):
    """TableFunc::operator() of the compiled reference at the points (x, y)."""
    data = np.ascontiguousarray(t.data, dtype=np.float64)
    hdr = np.ascontiguousarray(t.hdr, dtype=np.float64)
    x, y = np.broadcast_arrays(np.asarray(x, np.float64), np.asarray(y, np.float64))
    x, y = np.ascontiguousarray(x.ravel()), np.ascontiguousarray(y.ravel())
    out = np.zeros(len(x))
    ref_lib().ref_table_lookup(data.shape[0], data.shape[1], hdr.ctypes.data, data.ctypes.data, len(x),
                               x.ctypes.data, y.ctypes.data, out.ctypes.data)
    return out


def ref_p2c(W, eos=EOS_IDEAL):
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    U = np.zeros_like(W)
    ref_lib(eos).ref_p2c_n(W.shape[0], W.ctypes.data, U.ctypes.data)
    return W, U


def ref_c2p(U, W_init=None, eos=EOS_IDEAL):
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 5)
    W = np.zeros_like(U) if W_init is None else np.array(W_init, dtype=np.float64, order="C").reshape(-1, 5)
    codes = np.zeros(U.shape[0], np.int32)
    ref_lib(eos).ref_c2p_n(U.shape[0], U.ctypes.data, W.ctypes.data, codes.ctypes.data)
    return W, codes


def ref_pressure_fix(U, W, eos=EOS_IDEAL):
    U = np.array(U, dtype=np.float64, order="C").reshape(-1, 5)
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    codes = np.zeros(U.shape[0], np.int32)
    ref_lib(eos).ref_pressure_fix_n(U.shape[0], U.ctypes.data, W.ctypes.data, codes.ctypes.data)
    return U, W, codes


def ref_rotation(normals):
    normals = np.ascontiguousarray(normals, dtype=np.float64).reshape(-1, 3)
    T = np.zeros((normals.shape[0], 9))
    ref_lib().ref_rotation_n(normals.shape[0], normals.ctypes.data, T.ctypes.data)
    return T.reshape(-1, 3, 3)


def ref_flux(UL, UR, WL, WR, nS4, eos=EOS_IDEAL):
    arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(-1, 5) for a in (UL, UR, WL, WR)]
    nS4 = np.ascontiguousarray(nS4, dtype=np.float64).reshape(-1, 4)
    n = nS4.shape[0]
    F = np.zeros((n, 5))
    speed = np.zeros(n)
    ref_lib(eos).ref_flux_n(n, *[a.ctypes.data for a in arrs], nS4.ctypes.data, F.ctypes.data, speed.ctypes.data)
    return F, speed
