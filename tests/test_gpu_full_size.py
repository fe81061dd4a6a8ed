"""The benchmark configurations at their full sizes, checked through properties
that do not depend on the size (the oracle would take minutes there; it is
matched step for step at smaller sizes in test_gpu_step.py):

* config 1: the blast wave on 100^3 hexes x 6 = 6M tets, 100 steps;
* config 2: the W ~ 10 shear layer on 128^3 hexes x 6 = 12.6M tets with
  periodic faces, gamma = 4/3, 200 steps;
* config 3: the c2p surrogate inside config 2;
* config 4: the full form -- the blast wave on 159^3 hexes x 6 = 24.1M tets
  with the seeded FP32 radiation source, mixed precision, surrogate start,
  graph-captured, 100 steps.

Properties: mass and energy conservation (on the periodic box for every field;
on the blast's box while its outer cells are still untouched).  Both steps
add each face flux to its two cells with opposite signs and scale by tau/V
with V the FP64 volume, so the totals move only by rounding (bar 1e-13
relative).  With the radiation source (config 4) the momenta and energy take the source,
so only the mass balances: one step changes the total of V*D by tau times the
D flux through the boundary faces (their outer state is the cell's own).
Also: the mixed step against the FP64 step (per-field relative L1 <= 1e-5), no
c2p failure and the c2p residual |f(p)|/(E + p) <= 1e-10 in every cell, and the
FP64 step reproducing itself bitwise between eager steps and the captured graph.
"""
import numpy as np
import pytest

from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M
from radrhd_solver_amd import surrogate

pytestmark = pytest.mark.gpu

SEED = 20261202


def _run(m, W0, gamma, steps, precision, guess, model=None, graph=True):
    s = capi.Solver(m, gamma=gamma, cfl=0.4, precision=precision, guess=guess, collect_stats=True)
    try:
        if model is not None:
            s.set_mlp(model)
        s.upload_W(W0)
        s.sync()
        U0 = s.download()[1]
        if graph:
            s.run(steps)
        else:
            tau = 0.4 * m.h_min()
            for _ in range(steps):
                tau = 0.4 * m.h_min() / s.step(tau)
        s.sync()
        W, U = s.download()
        return U0, W, U, s.stats()
    finally:
        s.close()


def _rel_l1(a, b, V):
    return [float((np.abs(a[:, k] - b[:, k]) * V).sum() / (np.abs(b[:, k]) * V).sum()) for k in range(5)]


def _drift(U0, U, V):
    """|change of each field's total| over the field's total magnitude (start or end, the larger)."""
    t0 = np.array([(U0[:, k] * V).sum() for k in range(5)])
    t1 = np.array([(U[:, k] * V).sum() for k in range(5)])
    a = np.maximum(np.abs(U0).T @ V, np.abs(U).T @ V)
    return np.abs(t1 - t0) / a


def _clean(st, label):
    it = st["newton_iters"] / max(1, st["steps"])
    print(label, "c2p_fail", st["c2p_fail"], "fix_calls", st["fix_calls"], "max residual", st["max_residual"],
          "Newton iterations per step", round(it, 1), flush=True)
    assert sum(st["c2p_fail"]) == 0 and st["fix_calls"] == 0
    assert st["max_residual"] <= 1e-10


def test_config1_blast_6M():
    m = M.kuhn_box(100, 100, 100, seed=SEED)
    assert m.n_cells == 6_000_000
    W0 = M.init_state(m, M.INIT_BLAST)
    V = m.cells["V"]
    U0, W64, U64, st64 = _run(m, W0, 5.0 / 3.0, 100, capi.PREC_FP64, capi.GUESS_REFERENCE)
    _clean(st64, "fp64")
    _, _, U32, st32 = _run(m, W0, 5.0 / 3.0, 100, capi.PREC_MIXED, capi.GUESS_PREVIOUS)
    _clean(st32, "mixed")
    # the blast has not reached the outer cells (they move only by the FP64 closure
    # defect of their faces), so no mass or energy has crossed the boundary
    x = m.cells["center"]
    outer = np.abs(x - 0.5).max(axis=1) > 0.45
    assert np.abs(W64[outer, 1:4]).max() <= 1e-10
    assert np.abs(W64[outer, 4] / W0[outer, 4] - 1).max() <= 1e-10
    for label, U in (("fp64", U64), ("mixed", U32)):
        drift = _drift(U0, U, V)
        print(label, "relative drift of (D, mx, my, mz, E) totals", drift, flush=True)
        assert drift[0] <= 1e-13 and drift[4] <= 1e-13
    l1 = _rel_l1(U32, U64, V)
    print("mixed vs fp64 rel L1", l1, flush=True)
    assert max(l1) <= 1e-5


def test_config2_shear_12M():
    m = M.kuhn_box(128, 128, 128, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
    assert m.n_cells == 12_582_912
    W0 = M.init_state(m, M.INIT_SHEAR)
    V = m.cells["V"]
    U0, _, U64, st64 = _run(m, W0, 4.0 / 3.0, 200, capi.PREC_FP64, capi.GUESS_REFERENCE)
    _clean(st64, "fp64")
    _, _, U32, st32 = _run(m, W0, 4.0 / 3.0, 200, capi.PREC_MIXED, capi.GUESS_PREVIOUS)
    _clean(st32, "mixed")
    for label, U in (("fp64", U64), ("mixed", U32)):
        drift = _drift(U0, U, V)
        print(label, "relative drift of (D, mx, my, mz, E) totals", drift, flush=True)
        assert drift.max() <= 1e-13  # every face is interior: each flux enters two cells with opposite signs
    l1 = _rel_l1(U32, U64, V)
    print("mixed vs fp64 rel L1", l1, flush=True)
    assert max(l1) <= 1e-5


def test_config2_fp64_graph_equals_eager_12M():
    m = M.kuhn_box(128, 128, 128, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
    W0 = M.init_state(m, M.INIT_SHEAR)
    _, Wg, Ug, _ = _run(m, W0, 4.0 / 3.0, 20, capi.PREC_FP64, capi.GUESS_REFERENCE, graph=True)
    _, We, Ue, _ = _run(m, W0, 4.0 / 3.0, 20, capi.PREC_FP64, capi.GUESS_REFERENCE, graph=False)
    assert np.array_equal(Ug, Ue) and np.array_equal(Wg, We)


def test_config3_surrogate_in_shear_12M():
    """Config 3: the surrogate trained for gamma = 4/3 by the library's kernels, inside the
    config-2 shear layer.  (1) The step with the surrogate start (RHLLC_GUESS_MLP) against
    FP64: same residual bar, no c2p failure, L1 <= 1e-5; its gate (tau moved by more than
    5% in a step) selects no cell in this smooth flow, so it runs as the previous-pressure
    start does.  (2) Newton iteration histograms of the three starts on the same states
    (the shear layer after 1, 100 and 199 steps, the states one step later, 1M cells
    sampled): the previous pressure, the surrogate and the reference start rule."""
    model, info = surrogate.train(4.0 / 3.0)
    print("surrogate", {k: info[k] for k in ("steps", "batch", "loss_first", "loss_last", "seconds")}, flush=True)
    m = M.kuhn_box(128, 128, 128, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
    W0 = M.init_state(m, M.INIT_SHEAR)
    V = m.cells["V"]
    _, _, U64, _ = _run(m, W0, 4.0 / 3.0, 200, capi.PREC_FP64, capi.GUESS_REFERENCE)
    res = {}
    for guess in (capi.GUESS_PREVIOUS, capi.GUESS_MLP):
        _, _, U, st = _run(m, W0, 4.0 / 3.0, 200, capi.PREC_MIXED, guess, model=model if guess == capi.GUESS_MLP else None)
        _clean(st, f"guess {guess}")
        hist = np.array(st["newton_hist"])
        print(f"guess {guess}: Newton iteration histogram (bins 0..)", hist[:12].tolist(), "fallback cells",
              st["fallback_cells"], "surrogate starts", st["mlp_starts"], flush=True)
        l1 = _rel_l1(U, U64, V)
        print(f"guess {guess}: rel L1 vs fp64", l1, flush=True)
        assert max(l1) <= 1e-5
        res[guess] = st
    assert res[capi.GUESS_MLP]["fallback_cells"] <= res[capi.GUESS_PREVIOUS]["fallback_cells"] + 0.001 * m.n_cells * 200
    # (2) the three starts on the same states
    rng = np.random.default_rng(SEED + 3)
    sample = rng.choice(m.n_cells, 1 << 20, replace=False)
    s = capi.Solver(m, gamma=4.0 / 3.0, cfl=0.4, precision=capi.PREC_MIXED, guess=capi.GUESS_PREVIOUS)
    s.upload_W(W0)
    done, tot = 0, {"prev": [], "mlp": [], "ref": []}
    for k in (1, 100, 199):
        s.run(k - done, graph=False)
        s.sync()
        Wk, _ = s.download()
        s.step(s.stats()["tau"])
        done = k + 1
        U1 = s.download()[1][sample]
        starts = {"prev": Wk[sample, 4].astype(np.float32).astype(np.float64),
                  "mlp": capi.mlp_eval(model, U1), "ref": None}
        for name, g in starts.items():
            _, codes, it = capi.eval_c2p(U1, gamma=4.0 / 3.0, guess=g, precision=capi.PREC_MIXED)
            assert (codes == 0).all(), name
            tot[name].append(it)
            print(f"step {k} start {name}: mean Newton iterations {it.mean():.3f}, histogram",
                  np.bincount(it, minlength=8)[:8].tolist(), flush=True)
    s.close()
    mean = {k: float(np.concatenate(v).mean()) for k, v in tot.items()}
    print("mean Newton iterations over the probed states:", mean)
    assert mean["mlp"] < mean["ref"]  # the surrogate is the better cold start


def _mass_balance(s, m, V, tau):
    """One eager step: |change of total V*D + tau * boundary D flux| over the total."""
    _, U0 = s.download()
    s.step(tau)
    s.sync()
    _, U1 = s.download()
    F = s.face_flux()
    bnd = M.face_id_r(m.faces) < 0
    M0, M1 = (U0[:, 0] * V).sum(), (U1[:, 0] * V).sum()
    return abs(M1 - M0 + tau * F[bnd, 0].sum()) / M0, (M1 - M0) / M0


def test_config4_full_form_24M():
    """Config 4: 24.1M tets, blast wave + seeded radiation source, mixed precision with the
    surrogate start (the stored gamma = 5/3 model), graph-captured, 100 steps with the
    statistics on: no c2p failure, no pressure fix, residual <= 1e-10 in every cell, per-field
    relative L1 <= 1e-5 against the FP64 step under the same source, and the mass balance of
    one more step against the boundary flux (FP64 to rounding, mixed to its FP32 geometry)."""
    m = M.kuhn_box(159, 159, 159, seed=SEED)
    assert m.n_cells == 24_118_074
    W0 = M.init_state(m, M.INIT_BLAST)
    rad = M.init_radiation(m.n_cells, seed=SEED)
    V = m.cells["V"]
    model, info = surrogate.load_or_train(5.0 / 3.0)
    print("surrogate", info.get("source"), flush=True)
    h = m.h_min()
    tau = 0.4 * h
    out = {}
    for label, prec, guess in (("fp64", capi.PREC_FP64, capi.GUESS_REFERENCE),
                               ("mixed+mlp", capi.PREC_MIXED, capi.GUESS_MLP)):
        s = capi.Solver(m, gamma=5.0 / 3.0, cfl=0.4, h_min=h, precision=prec, guess=guess, collect_stats=True)
        try:
            s.upload_W(W0)
            s.set_radiation(*rad)
            if guess == capi.GUESS_MLP:
                s.set_mlp(model)
            s.run(100, graph=True)
            s.sync()
            st = s.stats()
            _clean(st, label)
            assert st["steps"] == 100
            print(label, "surrogate starts", st["mlp_starts"], "fallback cells", st["fallback_cells"], flush=True)
            out[label] = s.download()[1]
            bal, dm = _mass_balance(s, m, V, tau)
            print(label, "mass change of one step", dm, "imbalance vs boundary flux", bal, flush=True)
            assert bal <= 1e-13
        finally:
            s.close()
    l1 = _rel_l1(out["mixed+mlp"], out["fp64"], V)
    print("mixed+mlp vs fp64 rel L1 after 100 steps", l1, flush=True)
    assert max(l1) <= 1e-5
