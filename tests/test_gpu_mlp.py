"""GPU checks of the c2p surrogate MLP: MFMA lane layout, gradients and
forward pass against the CPU restatement (oracle/mlp.py), training, and the
surrogate-started c2p inside the mixed step."""
import numpy as np
import pytest

import oracle as O
from oracle import mlp as R
from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

pytestmark = pytest.mark.gpu
SEED = 20261202


def test_mfma_16x16x32_bf16_lane_layout():
    """The K = 32 BF16 MFMA as the MLP kernels feed it: A[i][3] = i + 1, A[i][23] = 1/2,
    B[3][j] = 16 (j + 1), B[23][j] = 2 (units 3 and 23 sit in different halves of the
    lanes' eight elements), so C[i][j] = 16 (i + 1)(j + 1) + 1 in lane j + 16 (i / 4)."""
    out = capi.mlp_selftest()
    lane = np.arange(64)
    g, c = lane // 16, lane % 16
    for r in range(4):
        # C[i][j] = 16 (i + 1)(j + 1) + 1 with i = 4 g + r, j = c
        assert np.array_equal(out[:, r], 16.0 * (4 * g + r + 1) * (c + 1) + 1.0), r


def _model(rng):
    m = np.zeros(R.N_MODEL, np.float32)
    m[: R.N_PARAMS] = rng.normal(scale=0.25, size=R.N_PARAMS)
    m[R.N_PARAMS:] = [0.5, 1.0, -2.0, 2.0, 1.5, 3.0, 0.5, 0.4, 5.0 / 3.0]
    return m


def test_gradient_matches_cpu_restatement():
    rng = np.random.default_rng(SEED)
    model = _model(rng)
    n = 1024
    X = rng.normal(size=(n, 3)).astype(np.float32)
    y = rng.normal(size=n).astype(np.float32)
    ok = (rng.uniform(size=n) > 0.05).astype(np.float32)
    g, loss = capi.mlp_grad(model, X, y, ok)
    lo, go = R.loss_and_grad(model[: R.N_PARAMS], X.astype(np.float64), y, ok, bf=True)
    assert abs(loss - lo) <= 2e-3 * lo
    for a, b, name in ((R.OFF_W1, R.OFF_B1, "W1"), (R.OFF_B1, R.OFF_W2, "b1"), (R.OFF_W2, R.OFF_B2, "W2"),
                       (R.OFF_B2, R.OFF_W3, "b2"), (R.OFF_W3, R.OFF_B3, "W3"), (R.OFF_B3, R.N_PARAMS, "b3")):
        gg, oo = g[a:b].astype(np.float64), go[a:b]
        rel = np.linalg.norm(gg - oo) / np.linalg.norm(oo)
        print(name, "rel L2", rel)
        assert rel < 2e-2, name
        if b - a > 1:
            assert gg @ oo / (np.linalg.norm(gg) * np.linalg.norm(oo)) > 0.9995, name


def test_forward_matches_cpu_restatement():
    rng = np.random.default_rng(SEED + 1)
    model = _model(rng)
    W = np.column_stack([10 ** rng.uniform(-2, 1, 3000), rng.uniform(-0.5, 0.5, (3000, 3)), 10 ** rng.uniform(-2, 2, 3000)])
    _, U = O.p2c(W)
    pg = capi.mlp_eval(model, U)
    po = R.predict_p(model, U, bf=True)
    rel = np.abs(np.log(pg) - np.log(po))
    assert np.median(rel) < 1e-4 and rel.max() < 5e-2


@pytest.fixture(scope="module")
def trained():
    model, hist = capi.mlp_train(steps=1500, batch=16384, lr=3e-3, seed=SEED)
    return model, hist


def test_training_reduces_loss_and_predicts_pressure(trained):
    model, hist = trained
    print("loss first/last", hist[:5].mean(), hist[-20:].mean())
    assert hist[-20:].mean() < 0.05 * hist[:5].mean()
    rng = np.random.default_rng(SEED + 7)
    n = 20000
    rho = 10 ** rng.uniform(-4, 2, n)
    p = rho * 10 ** rng.uniform(-6, 2, n)
    lor = np.exp(rng.uniform(0, np.log(50), n))
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    W = np.column_stack([rho, d * np.sqrt(1 - 1 / lor**2)[:, None], p])
    _, U = O.p2c(W)
    pg = capi.mlp_eval(model, U)
    rel = np.abs(pg / p - 1)
    print("surrogate |p/p_true - 1|: median", np.median(rel), "p90", np.percentile(rel, 90))
    assert np.median(rel) < 0.2
    # Newton from the surrogate start converges everywhere the reference start does
    Wg, cg, ig = capi.eval_c2p(U, guess=pg)
    Wr, cr, ir = capi.eval_c2p(U)
    assert (cg == 0).all() and (cr == 0).all()
    print("mean Newton iterations: surrogate", ig.mean(), "reference start", ir.mean())
    assert ig.mean() < ir.mean()
    assert np.allclose(Wg[:, 4], Wr[:, 4], rtol=1e-9)


def test_mixed_step_with_surrogate_start(trained):
    model, _ = trained
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    res = {}
    for guess in (capi.GUESS_PREVIOUS, capi.GUESS_MLP):
        s = capi.Solver(m, cfl=0.4, precision=capi.PREC_MIXED, guess=guess, collect_stats=True)
        if guess == capi.GUESS_MLP:
            s.set_mlp(model)
        s.upload_W(W0)
        s.run(40)
        s.sync()
        res[guess] = (s.download()[1], s.stats())
        s.close()
    Up, sp = res[capi.GUESS_PREVIOUS]
    Um, sm = res[capi.GUESS_MLP]
    print("prev-p stats", sp["newton_iters"], sp["fallback_cells"], "mlp stats", sm["newton_iters"],
          sm["fallback_cells"], "surrogate starts", sm["mlp_starts"])
    assert sm["mlp_starts"] > 0  # the blast front changes tau by more than the 5% gate
    assert sm["fix_fail"] == 0 and sum(sm["c2p_fail"]) == 0
    # both starts converge to the same Newton tolerance, so the states agree far below the mixed error
    scale = np.abs(Up).max(axis=0)
    assert (np.abs(Um - Up) <= 1e-9 * scale).all()


def test_surrogate_start_gains_on_the_cells_it_starts():
    """Newton iterations of the mixed c2p on exactly the cells the step defers to the
    surrogate (tau = E - D changed by more than 5%), on identical states: from the
    surrogate's pressure against the previous step's FP32 pressure (the start the step
    would otherwise use).  A 32^3-hex blast without radiation (the gate sees the hydro
    update, which is then the whole update); the states before and after one step are
    downloaded and the gate and both starts are evaluated on them."""
    from radrhd_solver_amd import surrogate

    model, _ = surrogate.load_or_train(5.0 / 3.0)
    m = M.kuhn_box(32, 32, 32, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    s = capi.Solver(m, cfl=0.4, precision=capi.PREC_MIXED, guess=capi.GUESS_PREVIOUS)
    s.upload_W(W0)
    total = {"prev": 0, "mlp": 0, "cells": 0}
    done = 0
    for k in (1, 5, 20, 40):
        s.run(k - done, graph=False)
        s.sync()
        Wk, Uk = s.download()
        s.step(s.stats()["tau"])
        done = k + 1
        _, U1 = s.download()
        t0, t1 = Uk[:, 4] - Uk[:, 0], U1[:, 4] - U1[:, 0]
        sel = np.abs(t1 - t0) > 0.05 * np.abs(t1)
        U = U1[sel]
        p_prev = Wk[sel, 4].astype(np.float32).astype(np.float64)
        _, cp, ip = capi.eval_c2p(U, guess=p_prev, precision=capi.PREC_MIXED)
        _, cm, im = capi.eval_c2p(U, guess=capi.mlp_eval(model, U), precision=capi.PREC_MIXED)
        assert (cp == 0).all() and (cm == 0).all()
        print(f"step {k}: {sel.sum()} deferred cells, mean Newton iterations previous-pressure start "
              f"{ip.mean():.3f}, surrogate start {im.mean():.3f}", flush=True)
        assert sel.sum() > 100
        total["prev"] += ip.sum()
        total["mlp"] += im.sum()
        total["cells"] += sel.sum()
    s.close()
    prev, mlp = total["prev"] / total["cells"], total["mlp"] / total["cells"]
    print("over all probed steps: previous pressure", prev, "surrogate", mlp)
    assert mlp <= prev - 0.5
