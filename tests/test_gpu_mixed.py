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
    """Mixed face flux (FP32 cell records and geometry, FP64 fast arithmetic, packed FP32 output) vs the
    FP64 flux of the unrounded states: 2e-5 of scale at worst, 1e-7 typical.  The worst cases are set by
    rounding the inputs to FP32 (states up to |v| = 0.9, where the flux's sensitivity to v grows like W^2),
    not by the arithmetic."""
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
    assert err.max() < 2e-5 and np.median(err) < 1e-7
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


@pytest.mark.parametrize("case", ["blast", "shear"])
def test_mixed_c2p_residual_and_iterations(case):
    """Every cell of the mixed step ends at |f(p)|/(E + p) <= 1e-10 (the residual the FP64 step
    reaches), with no c2p failure, on the blast wave and on a W ~ 10 periodic shear layer."""
    if case == "blast":
        m = M.kuhn_box(24, 24, 24, seed=SEED)
        W0, gamma, steps = M.init_state(m, M.INIT_BLAST), 5.0 / 3.0, 40
    else:
        m = M.kuhn_box(24, 24, 24, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
        W0, gamma, steps = M.init_state(m, M.INIT_SHEAR), 4.0 / 3.0, 40
    out = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = capi.Solver(m, gamma=gamma, cfl=0.4, precision=prec, guess=guess, collect_stats=True)
        s.upload_W(W0)
        s.run(steps)
        s.sync()
        out[prec] = s.stats()
        s.close()
    for prec, st in out.items():
        it = st["newton_iters"] / (m.n_cells * steps)
        print(case, "precision", prec, "mean Newton iterations", round(it, 3), "max residual", st["max_residual"])
        assert sum(st["c2p_fail"]) == 0 and st["fix_calls"] == 0
        assert st["max_residual"] <= 1e-10


@pytest.mark.parametrize("eos", [capi.EOS_IDEAL, capi.EOS_TAUB], ids=["ideal", "taub"])
@pytest.mark.parametrize("W", [(1.0, 0.0, 0.0, 0.0, 1.0), (1.0, 0.5, 0.3, -0.2, 2.0)], ids=["rest", "moving"])
def test_uniform_state_stays_uniform(W, eos):
    """A uniform state on a jittered periodic box: every face flux is the same physical flux, so a
    cell changes only by that flux through the closure defect of its four face vectors.  The mixed
    step's vectors (FP32 plus 7-bit residuals) close to 2^-30 of their size: after 50 steps every
    primitive stays within 1e-8 of the initial one (measured 3e-9; FP32 vectors alone, 2^-24: ~1e-6)."""
    m = M.kuhn_box(16, 16, 16, periodic=7, seed=SEED)
    W0 = M.init_state(m, M.INIT_UNIFORM, params=list(W))
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = capi.Solver(m, cfl=0.4, precision=prec, guess=guess, eos=eos)
        try:
            s.upload_W(W0)
            s.run(50)
            s.sync()
            Wg = s.download()[0]
        finally:
            s.close()
        dev = np.abs(Wg - W0).max(axis=0)
        print("precision", prec, "eos", eos, "max deviation of (rho, v, p) after 50 steps", dev, flush=True)
        assert dev.max() <= 1e-8
