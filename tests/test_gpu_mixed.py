"""GPU checks of the mixed FP32-flux / FP64-c2p path against the FP64 oracle.

The mixed path is not bitwise; these tests bound its error per face, per step
and over a run.  Tolerances are written in each test.
"""
import numpy as np
import pytest

import oracle as O
from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

pytestmark = pytest.mark.gpu
SEED = 20261202


def _prims(rng, n, vmax=0.9):
    r = 10 ** rng.uniform(-1, 1, n)
    p = 10 ** rng.uniform(-1, 2, n)
    sp = vmax * rng.uniform(0, 1, n)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.column_stack([r, d * sp[:, None], p])


def test_eval_flux_fp32_close_to_fp64():
    """Mixed face flux (FP32 inputs/geometry/output, FP64 fast arithmetic) vs the FP64 flux: 1e-5 of scale."""
    rng = np.random.default_rng(SEED)
    n = 20000
    WL, WR = _prims(rng, n), _prims(rng, n)
    WR[:2000] = WL[:2000]  # identical states take the physical flux exactly
    _, UL = O.p2c(WL)
    _, UR = O.p2c(WR)
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    nS = np.column_stack([nn, rng.uniform(0.5, 2, n)])
    Fo, so = O.flux(UL, UR, WL, WR, nS)
    Fg, sg = capi.eval_flux(UL, UR, WL, WR, nS, precision=capi.PREC_MIXED)
    scale = np.abs(Fo).max(axis=1) + np.abs(WL[:, 4]) * nS[:, 3]
    err = np.abs(Fg - Fo).max(axis=1) / scale
    print("fp32 flux rel err: max", err.max(), "median", np.median(err))
    assert err.max() < 1e-5
    assert np.abs(sg - so).max() < 1e-6


def test_mixed_first_step_face_flux_and_state():
    """One step from the same state: face fluxes within 1e-5 of scale, U within 1e-6 relative L1."""
    m = M.kuhn_box(8, 8, 8, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.5, 0.5, 0.5, 0.3, 1.0, 1000.0, 1.0])
    h = m.h_min()
    st = O.State(m, W0)
    st.run(0.4, h, 0)
    tau = 0.4 * h
    st.step(tau)
    s = capi.Solver(m, cfl=0.4, h_min=h, precision=capi.PREC_MIXED, guess=capi.GUESS_PREVIOUS, reorder=True)
    s.upload_W(W0)
    s.step(tau)
    Fg = s.face_flux()
    W, U = s.download()
    s.close()
    Fo = st.face_flux
    scale = np.abs(Fo).max(axis=1) + W0[M.face_id_l(m.faces), 4] * m.faces["S"]
    ferr = np.abs(Fg - Fo).max(axis=1) / scale
    print("first-step face flux rel err: max", ferr.max(), "worst face", int(ferr.argmax()))
    V = m.cells["V"]
    l1 = [float((np.abs(U[:, k] - st.U[:, k]) * V).sum() / max((np.abs(st.U[:, k]) * V).sum(), 1e-300))
          for k in range(5)]
    print("first-step U rel L1", l1)
    assert ferr.max() < 1e-5
    assert max(l1) < 1e-6
