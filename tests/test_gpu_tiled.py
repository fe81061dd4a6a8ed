"""The tiled step (RHLLC_SCHED_TILED) and the handle's error reporting on the GPU.

The tiled step computes a tile's faces into LDS and updates its cells from
there, with the gathered state double-buffered.  On the FP64 path it must give
the reference's bits (the oracle restates rhllc_3d_stab.cpp:10-65) whatever
the mix of graph replays (two steps each) and eager steps; the face fluxes it
reports are recomputed from the last step's input buffer.
"""
import numpy as np
import pytest

import oracle as O
from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

pytestmark = pytest.mark.gpu

SEED = 20261202


def _l1(a, b, V):
    return [float((np.abs(a[:, k] - b[:, k]) * V).sum() / (np.abs(b[:, k]) * V).sum()) for k in range(5)]


def test_tiled_buffer_parity_across_graph_and_eager_steps():
    """7 graph steps (three pair replays and one eager step), 3 eager, 5 graph, one
    host-tau step: 16 steps bitwise the oracle, stats and face fluxes included."""
    m = M.kuhn_box(20, 12, 10, seed=SEED)
    print("mesh", m.n_cells, m.n_faces, flush=True)
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.5, 0.5, 0.5, 0.2, 1.0, 100.0, 1.0])
    h = m.h_min()
    st = O.State(m, W0)
    t_o, _, speeds = st.run(0.4, h, 15)
    s = capi.Solver(m, cfl=0.4, h_min=h)
    s.upload_W(W0)
    s.run(7, graph=True)
    s.run(3, graph=False)
    s.run(5, graph=True)
    s.sync()
    W, U = s.download()
    stats = s.stats()
    assert np.array_equal(st.U, U) and np.array_equal(st.W, W)
    assert stats["steps"] == 15 and stats["t"] == t_o and stats["max_speed"] == speeds[-1]
    tau = 0.4 * h / speeds[-1]
    assert s.step(tau) == st.step(tau)
    W, U = s.download()
    assert np.array_equal(st.U, U) and np.array_equal(st.W, W)
    assert np.array_equal(s.face_flux(), st.face_flux)
    s.close()


def test_tiled_and_two_kernel_face_fluxes_agree():
    """The face fluxes of one step: tiled (recomputed from the input buffer) and two-kernel
    (the stored array) are the same bits on the FP64 path; on the mixed path the tiled
    fluxes are the unrounded FP64 values the step used."""
    m = M.kuhn_box(10, 10, 10, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    F = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        for sched in (capi.SCHED_TILED, capi.SCHED_TWO_KERNEL):
            s = capi.Solver(m, cfl=0.4, precision=prec, guess=guess, schedule=sched)
            s.upload_W(W0)
            assert not s.face_flux().any()  # no step yet
            s.run(3)
            s.sync()
            F[prec, sched] = s.face_flux()
            s.close()
    assert np.array_equal(F[0, 0], F[0, 1])
    scale = np.abs(F[0, 0]).max(axis=0)
    # mixed: FP32 states and geometry in both; the two-kernel record rounds the flux to 2^-31
    assert (np.abs(F[1, 0] - F[0, 0]).max(axis=0) <= 1e-5 * scale).all()
    assert (np.abs(F[1, 0] - F[1, 1]).max(axis=0) <= 1e-6 * scale).all()


@pytest.mark.parametrize("guess", [capi.GUESS_PREVIOUS, capi.GUESS_MLP])
def test_tiled_mixed_matches_fp64_and_two_kernel(guess):
    """Mixed step with the radiation source under both schedules, 40 steps of a 24^3-hex
    blast: per-field L1 against FP64 <= 1e-5, no c2p failure, residual <= 1e-10."""
    from radrhd_solver_amd import surrogate

    m = M.kuhn_box(24, 24, 24, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    rad = M.init_radiation(m.n_cells, seed=SEED)
    V = m.cells["V"]
    model = surrogate.load_or_train(5.0 / 3.0)[0] if guess == capi.GUESS_MLP else None
    out = {}
    for prec, g, sched in ((capi.PREC_FP64, capi.GUESS_REFERENCE, capi.SCHED_TWO_KERNEL),
                           (capi.PREC_MIXED, guess, capi.SCHED_TILED),
                           (capi.PREC_MIXED, guess, capi.SCHED_TWO_KERNEL)):
        s = capi.Solver(m, cfl=0.4, precision=prec, guess=g, schedule=sched, collect_stats=True)
        s.upload_W(W0)
        s.set_radiation(*rad)
        if g == capi.GUESS_MLP:
            s.set_mlp(model)
        s.run(40)
        s.sync()
        out[prec, sched] = (s.download()[1], s.stats())
        s.close()
    U64 = out[capi.PREC_FP64, capi.SCHED_TWO_KERNEL][0]
    for sched in (capi.SCHED_TILED, capi.SCHED_TWO_KERNEL):
        U, st = out[capi.PREC_MIXED, sched]
        l1 = _l1(U, U64, V)
        print("schedule", sched, "mixed vs fp64 L1", l1, "residual", st["max_residual"], flush=True)
        assert max(l1) <= 1e-5
        assert sum(st["c2p_fail"]) == 0 and st["fix_fail"] == 0
        assert st["max_residual"] <= 1e-10
        if guess == capi.GUESS_MLP:
            assert st["mlp_starts"] > 0


def test_upload_after_run_without_sync_waits_for_the_stream():
    """rhllc_run is asynchronous; an upload right after it must not race the queued steps."""
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    W1 = M.init_state(m, M.INIT_BLAST, params=[0.3, 0.6, 0.5, 0.15, 2.0, 50.0, 1.0])
    ref = capi.Solver(m, cfl=0.4)
    ref.upload_W(W1)
    ref.run(4)
    ref.sync()
    Wr, Ur = ref.download()
    ref.close()
    s = capi.Solver(m, cfl=0.4)
    s.upload_W(W0)
    s.run(50, graph=True)  # queued, not waited for
    s.upload_W(W1)
    s.run(4)
    s.sync()
    W, U = s.download()
    s.close()
    assert np.array_equal(U, Ur) and np.array_equal(W, Wr)


@pytest.mark.parametrize("schedule", [capi.SCHED_TILED, capi.SCHED_TWO_KERNEL])
def test_sync_reports_a_failed_pressure_fix(schedule):
    """A cell whose c2p and PhysPressureFix both fail (a NaN conserved state: c2p returns
    code 4, the secant never meets its tolerance): the reference aborts
    (rhllc_3d_stab.cpp:49-52); rhllc_run + rhllc_sync report RHLLC_ERR_C2P once."""
    m = M.kuhn_box(4, 4, 4, seed=SEED)
    W = M.init_state(m, M.INIT_BLAST)
    _, U = capi.eval_p2c(W)
    U[5] = np.nan
    s = capi.Solver(m, cfl=0.4, schedule=schedule)
    s.upload(W, U)
    s.run(1, graph=True)
    with pytest.raises(capi.RhllcError, match="PhysPressureFix"):
        s.sync()
    assert s.stats()["fix_fail"] >= 1
    s.sync()  # reported once
    s.close()


def test_add_rad_flux_keeps_the_fused_source_fields(tmp_path):
    """rhllc_add_rad_flux has its own moments; a later fused step still uses those of
    rhllc_set_radiation (the captured graph holds them)."""
    from rad_cases import physical_blast, table_path

    tab = M.read_table(table_path(tmp_path))
    m = M.kuhn_box(8, 8, 8, seed=SEED)
    W0, Ur, Fr = physical_blast(m)
    res = []
    for extra in (False, True):
        s = capi.Solver(m, cfl=0.4)
        s.set_opacity_table(tab)
        s.upload_W(W0)
        s.set_radiation(Ur, Fr)
        s.run(2)
        s.sync()
        if extra:  # a separate source with other moments and tau = 0 leaves the state as it is
            s.add_rad_flux(0.0, Ur * 0.5, Fr * 0.25)
        s.run(3)
        s.sync()
        res.append(s.download()[1])
        s.close()
    assert np.array_equal(res[0], res[1])
