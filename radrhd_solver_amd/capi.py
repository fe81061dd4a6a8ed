"""ctypes declarations of ``include/rhllc_hip.h`` and a thin Python handle.

This mirrors what a maintainer of the reference would bind; the product path
is the native library -- there is no Python or CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

PREC_FP64, PREC_MIXED = 0, 1
GUESS_REFERENCE, GUESS_PREVIOUS, GUESS_MLP = 0, 1, 2
EOS_IDEAL, EOS_TAUB = 0, 1
SCHED_TILED, SCHED_TWO_KERNEL = 0, 1
HIST_BINS = 41


MLP_PARAMS, MLP_MODEL = 4481, 4490


class MlpConfig(C.Structure):
    _fields_ = [
        ("device", C.c_int32),
        ("gamma", C.c_double),
        ("steps", C.c_int32),
        ("batch", C.c_int32),
        ("lr", C.c_double),
        ("seed", C.c_uint64),
        ("norm_samples", C.c_int32),
    ]


class Params(C.Structure):
    _fields_ = [
        ("gamma", C.c_double),
        ("cfl", C.c_double),
        ("h_min", C.c_double),
        ("precision", C.c_int32),
        ("c2p_guess", C.c_int32),
        ("reorder", C.c_int32),
        ("device", C.c_int32),
        ("collect_stats", C.c_int32),
        ("has_out_source", C.c_int32),
        ("out_source_W", C.c_double * 5),
        ("out_source_U", C.c_double * 5),
        ("eos", C.c_int32),
        ("schedule", C.c_int32),
        ("lock_wall", C.c_int32),
        ("has_inner_source", C.c_int32),
        ("inner_source_W", C.c_double * 5),
        ("inner_source_U", C.c_double * 5),
        ("reject_outer_surface", C.c_int32),
        ("out_source_copy", C.c_int32),
    ]


class Stats(C.Structure):
    _fields_ = [
        ("steps", C.c_int64),
        ("c2p_fail", C.c_int64 * 6),
        ("fix_calls", C.c_int64),
        ("fix_fail", C.c_int64),
        ("fallback_cells", C.c_int64),
        ("mlp_starts", C.c_int64),
        ("newton_iters", C.c_int64),
        ("newton_hist", C.c_int64 * HIST_BINS),
        ("max_speed", C.c_double),
        ("tau", C.c_double),
        ("t", C.c_double),
        ("max_residual", C.c_double),
    ]

    def as_dict(self):
        return {
            "steps": self.steps,
            "c2p_fail": list(self.c2p_fail)[1:],
            "fix_calls": self.fix_calls,
            "fix_fail": self.fix_fail,
            "fallback_cells": self.fallback_cells,
            "mlp_starts": self.mlp_starts,
            "newton_iters": self.newton_iters,
            "newton_hist": list(self.newton_hist),
            "max_speed": self.max_speed,
            "tau": self.tau,
            "t": self.t,
            "max_residual": self.max_residual,
        }


# Every symbol the header declares: name -> (restype, argtypes)
_vp, _i64, _i32, _dbl = C.c_void_p, C.c_int64, C.c_int32, C.c_double
SIGNATURES = {
    "rhllc_create": (C.c_int, [_vp, _i64, _vp, _i64, _vp, _vp]),
    "rhllc_destroy": (C.c_int, [_vp]),
    "rhllc_upload_state": (C.c_int, [_vp, _vp, _i64, _i64, _i64]),
    "rhllc_download_state": (C.c_int, [_vp, _vp, _i64, _i64, _i64]),
    "rhllc_step": (C.c_int, [_vp, _dbl, _vp]),
    "rhllc_run": (C.c_int, [_vp, _i32, _i32]),
    "rhllc_sync": (C.c_int, [_vp]),
    "rhllc_set_time": (C.c_int, [_vp, _dbl, _dbl]),
    "rhllc_set_collect_stats": (C.c_int, [_vp, _i32]),
    "rhllc_get_stats": (C.c_int, [_vp, _vp]),
    "rhllc_reset_stats": (C.c_int, [_vp]),
    "rhllc_download_face_flux": (C.c_int, [_vp, _vp]),
    "rhllc_set_radiation": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "rhllc_set_radiation_emission": (C.c_int, [_vp, _i32]),
    "rhllc_add_rad_flux": (C.c_int, [_vp, _dbl, _vp, _vp]),
    "rhllc_add_rad_flux_tensor": (C.c_int, [_vp, _dbl, _vp, _vp, _vp]),
    "rhllc_set_radiation_tensor": (C.c_int, [_vp, _vp]),
    "rhllc_set_opacity_table": (C.c_int, [_vp, _i32, _i32, _vp, _vp]),
    "rhllc_time_kernels": (C.c_int, [_vp, _i32, _vp]),
    "rhllc_info": (C.c_int, [_vp, _vp, _vp, _vp]),
    "rhllc_tile_layout": (C.c_int, [_i64, _vp, _i64, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_pack_face_vectors": (C.c_int, [_i64, _vp, _vp]),
    "rhllc_eval_flux": (C.c_int, [_i32, _i32, _dbl, _i64, _vp, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_eval_c2p": (C.c_int, [_i32, _i32, _dbl, _i64, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_eval_p2c": (C.c_int, [_i32, _dbl, _i64, _vp, _vp]),
    "rhllc_eval_pressure_fix": (C.c_int, [_i32, _dbl, _i64, _vp, _vp, _vp]),
    "rhllc_eval_rad_source": (C.c_int, [_i32, _i64, _vp, _vp, _vp, _vp, _vp, _i32, _vp, _vp]),
    "rhllc_eval_rad_source_opt": (C.c_int, [_i32, _i32, _vp, _vp, _i64, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_device_query": (C.c_int, [_i32, C.c_char_p, _i32, _vp, _vp]),
    "rhllc_copy_probe": (C.c_int, [_i32, _i64, _i32, _vp]),
    "rhllc_mlp_train": (C.c_int, [_vp, _vp, _vp]),
    "rhllc_mlp_eval": (C.c_int, [_i32, _vp, _i64, _vp, _vp]),
    "rhllc_mlp_grad": (C.c_int, [_i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_mlp_selftest": (C.c_int, [_i32, _vp]),
    "rhllc_mlp_grad_noalloc": (C.c_int, [_i32, _vp, _i32, _vp, _vp, _vp, _vp, _vp, _vp]),
    "rhllc_set_mlp": (C.c_int, [_vp, _vp]),
    "rhllc_last_error": (C.c_char_p, []),
}


def declare(lib: C.CDLL) -> C.CDLL:
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


class RhllcError(RuntimeError):
    pass


def _lib():
    from ._lib import load_hip_lib

    return load_hip_lib()


def check(rc: int) -> None:
    if rc != 0:
        raise RhllcError(f"rhllc error {rc}: {_lib().rhllc_last_error().decode()}")


def _ptr(a):
    return None if a is None else a.ctypes.data


class Solver:
    """Owns one device handle (``rhllc_solver``)."""

    def __init__(self, mesh, gamma=5.0 / 3.0, cfl=0.4, h_min=None, precision=PREC_FP64,
                 guess=GUESS_REFERENCE, reorder=True, device=0, collect_stats=False, out_source=None,
                 eos=EOS_IDEAL, schedule=SCHED_TILED, inner_source=None, lock_wall=False,
                 reject_outer_surface=False, out_source_copy=False):
        L = _lib()
        p = Params()
        p.gamma, p.cfl = gamma, cfl
        p.h_min = mesh.h_min() if h_min is None else h_min
        p.precision, p.c2p_guess, p.reorder, p.device = precision, guess, int(reorder), device
        p.collect_stats = int(collect_stats)
        p.eos = eos
        p.schedule = schedule
        if out_source is not None:
            p.has_out_source = 1
            for k in range(5):
                p.out_source_W[k] = out_source[0][k]
                p.out_source_U[k] = out_source[1][k]
        if inner_source is not None:
            p.has_inner_source = 1
            for k in range(5):
                p.inner_source_W[k] = inner_source[0][k]
                p.inner_source_U[k] = inner_source[1][k]
        p.lock_wall = int(bool(lock_wall))
        p.reject_outer_surface = int(bool(reject_outer_surface))
        p.out_source_copy = int(bool(out_source_copy))
        self.params = p
        self.n_cells, self.n_faces = mesh.n_cells, mesh.n_faces
        h = C.c_void_p()
        check(L.rhllc_create(C.byref(p), mesh.n_cells, mesh.cells.ctypes.data, mesh.n_faces,
                             mesh.faces.ctypes.data, C.byref(h)))
        self.h = h

    def close(self):
        if self.h:
            _lib().rhllc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_W(self, W):
        """Primitive state W[N,5]; U = p2c(W) on the device (rhllc::Init)."""
        W = np.ascontiguousarray(W, dtype=np.float64).reshape(self.n_cells, 5)
        check(_lib().rhllc_upload_state(self.h, W.ctypes.data, 40, 0, -1))

    def upload(self, W, U):
        rec = np.ascontiguousarray(np.concatenate([np.asarray(W, np.float64), np.asarray(U, np.float64)], axis=1))
        check(_lib().rhllc_upload_state(self.h, rec.ctypes.data, 80, 0, 40))

    def download(self):
        rec = np.zeros((self.n_cells, 10))
        check(_lib().rhllc_download_state(self.h, rec.ctypes.data, 80, 0, 40))
        return rec[:, :5].copy(), rec[:, 5:].copy()

    def step(self, tau: float) -> float:
        s = C.c_double()
        check(_lib().rhllc_step(self.h, tau, C.byref(s)))
        return s.value

    def run(self, nsteps: int, graph: bool = True):
        check(_lib().rhllc_run(self.h, nsteps, int(graph)))

    def sync(self):
        check(_lib().rhllc_sync(self.h))

    def set_time(self, tau: float, t: float = 0.0):
        check(_lib().rhllc_set_time(self.h, tau, t))

    def set_collect_stats(self, on: bool):
        check(_lib().rhllc_set_collect_stats(self.h, int(on)))

    def stats(self) -> dict:
        s = Stats()
        check(_lib().rhllc_get_stats(self.h, C.byref(s)))
        return s.as_dict()

    def reset_stats(self):
        check(_lib().rhllc_reset_stats(self.h))

    def face_flux(self):
        F = np.zeros((self.n_faces, 5))
        check(_lib().rhllc_download_face_flux(self.h, F.ctypes.data))
        return F

    def set_radiation(self, Ur, Fr, ka=None, ks=None):
        """Per-cell radiation moments (and opacities, which an opacity table replaces)."""
        if Ur is None:
            check(_lib().rhllc_set_radiation(self.h, None, None, None, None))
            return
        arrs = [None if a is None else np.ascontiguousarray(a, dtype=np.float32) for a in (Ur, Fr, ka, ks)]
        self._rad = arrs
        check(_lib().rhllc_set_radiation(self.h, *[None if a is None else a.ctypes.data for a in arrs]))

    def add_rad_flux(self, tau, Ur, Fr, Tr=None):
        """AddRadFlux + HllcConvToPhys with the table's opacities on the current state; with
        Tr (n x 3 x 3) the pressure-tensor branch (ISOTROPIC_INTENSITY 0)."""
        Ur = np.ascontiguousarray(Ur, dtype=np.float32)
        Fr = np.ascontiguousarray(Fr, dtype=np.float32)
        if Tr is None:
            check(_lib().rhllc_add_rad_flux(self.h, tau, Ur.ctypes.data, Fr.ctypes.data))
        else:
            Tr = np.ascontiguousarray(Tr, dtype=np.float32).reshape(self.n_cells, 9)
            check(_lib().rhllc_add_rad_flux_tensor(self.h, tau, Ur.ctypes.data, Fr.ctypes.data, Tr.ctypes.data))

    def set_radiation_tensor(self, Tr):
        """Per-cell radiation pressure tensor (n x 3 x 3) for the fused source; None: M1 closure."""
        if Tr is None:
            check(_lib().rhllc_set_radiation_tensor(self.h, None))
            return
        self._Tr = np.ascontiguousarray(Tr, dtype=np.float32).reshape(self.n_cells, 9)
        check(_lib().rhllc_set_radiation_tensor(self.h, self._Tr.ctypes.data))

    def set_radiation_emission(self, on):
        check(_lib().rhllc_set_radiation_emission(self.h, 1 if on else 0))

    def set_opacity_table(self, table):
        """table: (header[6], data[nx, ny]) as read_table returns, or None to remove it."""
        if table is None:
            check(_lib().rhllc_set_opacity_table(self.h, 0, 0, None, None))
            return
        hdr, data = table
        hdr = np.ascontiguousarray(hdr, dtype=np.float64).reshape(6)
        data = np.ascontiguousarray(data, dtype=np.float64)
        check(_lib().rhllc_set_opacity_table(self.h, data.shape[0], data.shape[1], hdr.ctypes.data,
                                             data.ctypes.data))

    def set_mlp(self, model):
        model = np.ascontiguousarray(model, dtype=np.float32)
        assert model.size == MLP_MODEL
        check(_lib().rhllc_set_mlp(self.h, model.ctypes.data))

    def time_kernels(self, nsteps: int):
        out = np.zeros(4)
        check(_lib().rhllc_time_kernels(self.h, nsteps, out.ctypes.data))
        return {"face_ms": out[0], "cell_ms": out[1], "dt_ms": out[2], "step_ms": out[3]}

    def algorithmic_bytes(self) -> int:
        b = C.c_int64()
        check(_lib().rhllc_info(self.h, None, None, C.byref(b)))
        return b.value


def tile_layout(mesh, reorder=True):
    """The tiled step's host layout of a mesh (no device): a dict of arrays."""
    L = _lib()
    nt, ne, ms = C.c_int32(), C.c_int64(), C.c_int32()
    args = (mesh.n_cells, mesh.cells.ctypes.data, mesh.n_faces, mesh.faces.ctypes.data, int(reorder))
    check(L.rhllc_tile_layout(*args, C.byref(nt), C.byref(ne), C.byref(ms), None, None, None, None, None))
    out = {
        "cell_new_to_old": np.zeros(mesh.n_cells, np.int32),
        "face_new_to_old": np.zeros(mesh.n_faces, np.int32),
        "tile_meta": np.zeros((nt.value, 4), np.int32),
        "tile_extra": np.zeros(max(ne.value, 1), np.int32),
        "cell_slots": np.zeros((mesh.n_cells, 2), np.uint32),
    }
    check(L.rhllc_tile_layout(*args, C.byref(nt), C.byref(ne), C.byref(ms),
                              *[out[k].ctypes.data for k in ("cell_new_to_old", "face_new_to_old", "tile_meta",
                                                              "tile_extra", "cell_slots")]))
    out["tile_extra"] = out["tile_extra"][:ne.value]
    out["max_slots"] = ms.value
    return out


def pack_face_vectors(Sn):
    """The mixed path's 16-byte form of face vectors S n (n x 3 float64) as rhllc_create stores
    them: an (n, 4) float32 array whose fourth column holds the residual bits (no device)."""
    Sn = np.ascontiguousarray(Sn, dtype=np.float64).reshape(-1, 3)
    out = np.zeros((Sn.shape[0], 4), np.float32)
    check(_lib().rhllc_pack_face_vectors(Sn.shape[0], Sn.ctypes.data, out.ctypes.data))
    return out


# ----------------------------------------------------------- pointwise evals
def eval_flux(UL, UR, WL, WR, nS4, gamma=5.0 / 3.0, precision=PREC_FP64, eos=EOS_IDEAL):
    a = [np.ascontiguousarray(x, dtype=np.float64).reshape(-1, 5) for x in (UL, UR, WL, WR)]
    nS4 = np.ascontiguousarray(nS4, dtype=np.float64).reshape(-1, 4)
    n = nS4.shape[0]
    F = np.zeros((n, 5))
    sp = np.zeros(n)
    check(_lib().rhllc_eval_flux(precision, eos, gamma, n, *[x.ctypes.data for x in a], nS4.ctypes.data,
                                 F.ctypes.data, sp.ctypes.data))
    return F, sp


def eval_c2p(U, gamma=5.0 / 3.0, guess=None, W_init=None, precision=PREC_FP64, eos=EOS_IDEAL):
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 5)
    n = U.shape[0]
    W = np.zeros((n, 5)) if W_init is None else np.array(W_init, dtype=np.float64, order="C").reshape(n, 5)
    g = None if guess is None else np.ascontiguousarray(guess, dtype=np.float64)
    codes = np.zeros(n, np.int32)
    iters = np.zeros(n, np.int32)
    check(_lib().rhllc_eval_c2p(precision, eos, gamma, n, U.ctypes.data, _ptr(g), W.ctypes.data, codes.ctypes.data,
                                iters.ctypes.data))
    return W, codes, iters


def eval_p2c(W, gamma=5.0 / 3.0, eos=EOS_IDEAL):
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    U = np.zeros_like(W)
    check(_lib().rhllc_eval_p2c(eos, gamma, W.shape[0], W.ctypes.data, U.ctypes.data))
    return W, U


def eval_pressure_fix(U, W, gamma=5.0 / 3.0, eos=EOS_IDEAL):
    U = np.array(U, dtype=np.float64, order="C").reshape(-1, 5)
    W = np.array(W, dtype=np.float64, order="C").reshape(-1, 5)
    codes = np.zeros(U.shape[0], np.int32)
    check(_lib().rhllc_eval_pressure_fix(eos, gamma, U.shape[0], U.ctypes.data, W.ctypes.data, codes.ctypes.data))
    return U, W, codes


def _tensor(Tr, n):
    return None if Tr is None else np.ascontiguousarray(Tr, dtype=np.float32).reshape(n, 9)


def eval_rad_source(W, Ur, Fr, ka, ks, emission=False, Tr=None, precision=PREC_FP64):
    W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1, 5)
    n = W.shape[0]
    Ur, ka, ks = (np.ascontiguousarray(x, dtype=np.float64).reshape(n) for x in (Ur, ka, ks))
    Fr = np.ascontiguousarray(Fr, dtype=np.float64).reshape(n, 3)
    T = _tensor(Tr, n)
    G = np.zeros((n, 4))
    check(_lib().rhllc_eval_rad_source(precision, n, W.ctypes.data, Ur.ctypes.data, Fr.ctypes.data, ka.ctypes.data,
                                       ks.ctypes.data, 1 if emission else 0, _ptr(T), G.ctypes.data))
    return G


def eval_rad_source_opt(table, W, Ur, Fr, Tr=None):
    """GetRadSourceOpt with the table's opacities: (G[n, 4], cell[n, 5] = (T, vel, lorenz, alpha, betta))."""
    hdr, data = table
    hdr = np.ascontiguousarray(hdr, dtype=np.float64).reshape(6)
    data = np.ascontiguousarray(data, dtype=np.float64)
    W = np.ascontiguousarray(W, dtype=np.float64).reshape(-1, 5)
    n = W.shape[0]
    Ur = np.ascontiguousarray(Ur, dtype=np.float64).reshape(n)
    Fr = np.ascontiguousarray(Fr, dtype=np.float64).reshape(n, 3)
    T = _tensor(Tr, n)
    G = np.zeros((n, 4))
    cell = np.zeros((n, 5))
    check(_lib().rhllc_eval_rad_source_opt(data.shape[0], data.shape[1], hdr.ctypes.data, data.ctypes.data, n,
                                           W.ctypes.data, Ur.ctypes.data, Fr.ctypes.data, _ptr(T), G.ctypes.data,
                                           cell.ctypes.data))
    return G, cell


def mlp_train(steps=20000, batch=65536, lr=2e-3, gamma=5.0 / 3.0, seed=20261202, norm_samples=1 << 20, device=0):
    """Trains the c2p surrogate on the device; returns (model[4490] float32, loss per step)."""
    cfg = MlpConfig(device, gamma, steps, batch, lr, seed, norm_samples)
    model = np.zeros(MLP_MODEL, np.float32)
    hist = np.zeros(steps, np.float32)
    check(_lib().rhllc_mlp_train(C.byref(cfg), model.ctypes.data, hist.ctypes.data))
    return model, hist


def mlp_eval(model, U, device=0):
    model = np.ascontiguousarray(model, dtype=np.float32)
    U = np.ascontiguousarray(U, dtype=np.float64).reshape(-1, 5)
    p = np.zeros(U.shape[0])
    check(_lib().rhllc_mlp_eval(device, model.ctypes.data, U.shape[0], U.ctypes.data, p.ctypes.data))
    return p


def mlp_grad(model, X, y, ok, device=0):
    model = np.ascontiguousarray(model, dtype=np.float32)
    X = np.ascontiguousarray(X, dtype=np.float32).reshape(-1, 3)
    y = np.ascontiguousarray(y, dtype=np.float32)
    ok = np.ascontiguousarray(ok, dtype=np.float32)
    g = np.zeros(MLP_PARAMS, np.float32)
    loss = C.c_float()
    check(_lib().rhllc_mlp_grad(device, model.ctypes.data, X.shape[0], X.ctypes.data, y.ctypes.data, ok.ctypes.data,
                                g.ctypes.data, C.byref(loss)))
    return g, loss.value


def mlp_selftest(device=0):
    out = np.zeros(256, np.float32)
    check(_lib().rhllc_mlp_selftest(device, out.ctypes.data))
    return out.reshape(64, 4)


def copy_probe(nbytes: int = 1 << 30, reps: int = 20, device: int = 0) -> float:
    """Read+write GB/s of a plain copy kernel over nbytes."""
    g = C.c_double()
    check(_lib().rhllc_copy_probe(device, nbytes, reps, C.byref(g)))
    return g.value


def device_query(device: int = 0):
    name = C.create_string_buffer(128)
    cus = C.c_int32()
    gb = C.c_double()
    check(_lib().rhllc_device_query(device, name, 128, C.byref(cus), C.byref(gb)))
    return {"arch": name.value.decode(), "cus": cus.value, "hbm_gb": gb.value}
