"""The boundary branches of GEOMETRY_TYPE builds other than the default Cone
(solvers/rhllc/src/rhllc_utils.cpp:45-133): a wall for e_bound_lock (-2), a
fixed state for e_bound_inner_source (-4), the copy for e_bound_out_source
(-3).  The FP64 step against the oracle, bitwise, on both schedules; the mixed
step against FP64; a box closed by walls conserving mass and energy.
"""
import numpy as np
import pytest

import oracle as O
from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

pytestmark = pytest.mark.gpu

SEED = 20261202
W_INNER = np.array([0.1, 0.99, 0.0, 0.0, 0.01])  # the Cone_JET / Cylinder source state


def _mark(m, axis, side, code):
    r = M.face_id_r(m.faces)
    sel = (r < 0) & (np.sign(m.faces["n"][:, axis]) == side) & (np.abs(m.faces["n"][:, axis]) > 0.9)
    m.faces["w1"][sel] = (m.faces["w1"][sel] & np.uint32(0xFC000000)) | np.uint32(code & 0x3FFFFFF)
    assert sel.sum() > 0


def _mesh():
    """A jittered box whose x-min faces are walls, x-max faces the inner source, y-min faces
    out_source (copied), y-max outer_surface (copied); z faces free."""
    m = M.kuhn_box(10, 8, 8, seed=SEED)
    _mark(m, 0, -1, -2)
    _mark(m, 0, 1, -4)
    _mark(m, 1, -1, -3)
    _mark(m, 1, 1, -5)
    return m


def _inner_U(gamma):
    # rhllc::GetConvValue (rhllc_flux.cpp:7-16), ideal gas
    W = W_INNER
    G = 1.0 / np.sqrt(1 - (W[1] ** 2 + W[2] ** 2 + W[3] ** 2))
    h = 1 + gamma / (gamma - 1) * W[4] / W[0]
    dhGG = W[0] * h * G * G
    return np.array([G * W[0], dhGG * W[1], dhGG * W[2], dhGG * W[3], dhGG - W[4]])


def _solver(m, precision, guess, schedule, gamma=5.0 / 3.0):
    return capi.Solver(m, gamma=gamma, cfl=0.4, h_min=m.h_min(), precision=precision, guess=guess,
                       schedule=schedule, lock_wall=True, inner_source=(W_INNER, _inner_U(gamma)),
                       out_source_copy=True, collect_stats=True)


@pytest.mark.parametrize("schedule", [capi.SCHED_TILED, capi.SCHED_TWO_KERNEL])
def test_walls_and_sources_bitwise(schedule):
    m = _mesh()
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.3, 0.5, 0.5, 0.2, 1.0, 100.0, 1.0])
    st = O.State(m, W0)
    st.ph.lock_wall = 1
    st.ph.has_inner_source = 1
    U = _inner_U(5.0 / 3.0)
    for k in range(5):
        st.ph.inner_source_W[k], st.ph.inner_source_U[k] = W_INNER[k], U[k]
    steps = 40
    st.run(0.4, m.h_min(), steps)
    print("schedule", schedule, "steps", steps, flush=True)
    s = _solver(m, capi.PREC_FP64, capi.GUESS_REFERENCE, schedule)
    try:
        s.upload_W(W0)
        s.run(steps)
        s.sync()
        Wg, Ug = s.download()
        F = s.face_flux()
    finally:
        s.close()
    assert np.array_equal(Wg, st.W) and np.array_equal(Ug, st.U)
    assert np.array_equal(F, st.face_flux)
    r = M.face_id_r(m.faces)
    assert np.abs(F[r == -4]).max() > 0 and np.abs(st.W[:, 1]).max() > 0.1  # the sources and the blast act


def test_walls_and_sources_mixed_l1():
    m = _mesh()
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.3, 0.5, 0.5, 0.2, 1.0, 100.0, 1.0])
    V = m.cells["V"]
    out = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = _solver(m, prec, guess, capi.SCHED_TILED)
        try:
            s.upload_W(W0)
            s.run(40)
            s.sync()
            out[prec] = s.download()[1]
            st = s.stats()
        finally:
            s.close()
        assert sum(st["c2p_fail"]) == 0 and st["fix_calls"] == 0
    a, b = out[capi.PREC_MIXED], out[capi.PREC_FP64]
    l1 = [float((np.abs(a[:, k] - b[:, k]) * V).sum() / (np.abs(b[:, k]) * V).sum()) for k in range(5)]
    print("mixed vs fp64 rel L1 with walls and sources", l1, flush=True)
    assert max(l1) <= 1e-5


@pytest.mark.parametrize("eos", [capi.EOS_IDEAL, capi.EOS_TAUB], ids=["ideal", "taub"])
@pytest.mark.parametrize("precision", [capi.PREC_FP64, capi.PREC_MIXED])
def test_walled_box_conserves_mass_and_energy(precision, eos):
    """Every boundary face a wall: the blast moves no mass or energy out of the box."""
    m = M.kuhn_box(12, 12, 12, seed=SEED)
    for axis in range(3):
        for side in (-1, 1):
            _mark(m, axis, side, -2)
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.5, 0.5, 0.5, 0.25, 1.0, 100.0, 1.0])
    V = m.cells["V"]
    guess = capi.GUESS_REFERENCE if precision == capi.PREC_FP64 else capi.GUESS_PREVIOUS
    s = capi.Solver(m, cfl=0.4, precision=precision, guess=guess, lock_wall=True, eos=eos)
    try:
        s.upload_W(W0)
        s.sync()
        U0 = s.download()[1]
        s.run(80)
        s.sync()
        W, U = s.download()
    finally:
        s.close()
    assert np.abs(W[:, 1:4]).max() > 1e-2
    tot0, tot1 = (U0 * V[:, None]).sum(0), (U * V[:, None]).sum(0)
    scale = (np.abs(U) * V[:, None]).sum(0)
    drift = np.abs(tot1 - tot0) / scale
    print("precision", precision, "eos", eos, "relative drift of the totals", drift, flush=True)
    assert drift[0] <= 1e-13 and drift[4] <= 1e-13
