"""GPU parity of the HIP RHLLC step against the oracle (FP64: bitwise).

The oracle is the C restatement of the reference's Hllc3dStab
(solvers/rhllc/src/rhllc_3d_stab.cpp:10-65), itself pinned bit-for-bit to the
compiled reference in tests/test_oracle_vs_ref.py and by tests/golden.
"""
import numpy as np
import pytest

import oracle as O
from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

pytestmark = pytest.mark.gpu

SEED = 20261202


def random_prims(rng, n, vmax=0.999, rho=(-3, 2), p=(-4, 3)):
    r = 10 ** rng.uniform(*rho, n)
    pr = 10 ** rng.uniform(*p, n)
    speed = vmax * rng.uniform(0, 1, n) ** 0.5
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.column_stack([r, d * speed[:, None], pr])


def random_faces(rng, n):
    nn = rng.normal(size=(n, 3))
    nn /= np.linalg.norm(nn, axis=1)[:, None]
    nn[:8] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    return np.column_stack([nn, rng.uniform(0.01, 2.0, n)])


def test_device_is_gfx950():
    info = capi.device_query(0)
    assert info["arch"].startswith("gfx950"), info


@pytest.mark.parametrize("gamma", [5.0 / 3.0, 4.0 / 3.0])
def test_eval_flux_fp64_bitwise(gamma):
    rng = np.random.default_rng(SEED)
    n = 20000
    WL, WR = random_prims(rng, n), random_prims(rng, n)
    # supersonic left/right-going pairs exercise the lL >= 0 and lR <= 0 branches
    WL[:500, 1:4] = [0.99, 0, 0]
    WR[:500, 1:4] = [0.99, 0, 0]
    WL[500:1000, 1:4] = [-0.99, 0, 0]
    WR[500:1000, 1:4] = [-0.99, 0, 0]
    _, UL = O.p2c(WL, gamma)
    _, UR = O.p2c(WR, gamma)
    nS = random_faces(rng, n)
    nS[:1000, :3] = [1, 0, 0]
    Fo, so = O.flux(UL, UR, WL, WR, nS, gamma)
    Fg, sg = capi.eval_flux(UL, UR, WL, WR, nS, gamma)
    assert np.array_equal(Fo, Fg), np.argwhere(Fo != Fg)[:5]
    assert np.array_equal(so, sg)


def test_eval_c2p_p2c_fix_bitwise():
    rng = np.random.default_rng(SEED + 1)
    n = 20000
    W = random_prims(rng, n)
    Wc_o, U_o = O.p2c(W)
    Wc_g, U_g = capi.eval_p2c(W)
    assert np.array_equal(U_o, U_g) and np.array_equal(Wc_o, Wc_g)
    # perturb energies so some states fail (Q < 0) and go to the fix
    U = U_o.copy()
    U[:2000, 4] = np.sqrt(U[:2000, 0] ** 2 + (U[:2000, 1:4] ** 2).sum(1)) * (1 - 1e-3)
    W0 = np.full((n, 5), -7.0)
    Wo, co, io = O.c2p(U, W_init=W0)
    Wg, cg, ig = capi.eval_c2p(U, W_init=W0)
    assert np.array_equal(co, cg)
    assert np.array_equal(io, ig)
    assert np.array_equal(Wo, Wg)
    assert (co != 0).sum() >= 2000
    bad = co != 0
    Uo2, Wo2, fo = O.pressure_fix(U[bad], Wo[bad])
    Ug2, Wg2, fg = capi.eval_pressure_fix(U[bad], Wg[bad])
    assert np.array_equal(fo, fg) and np.array_equal(Uo2, Ug2) and np.array_equal(Wo2, Wg2)


def _run_both(mesh, W0, cfl, steps, gamma=5.0 / 3.0, graph=True, reorder=True, eos=capi.EOS_IDEAL,
              schedule=capi.SCHED_TILED):
    h = mesh.h_min()
    st = O.State(mesh, W0, gamma, eos=eos)
    t_o, taus_o, speeds_o = st.run(cfl, h, steps)
    s = capi.Solver(mesh, gamma=gamma, cfl=cfl, h_min=h, reorder=reorder, eos=eos, schedule=schedule)
    s.upload_W(W0)
    s.run(steps, graph=graph)
    s.sync()
    Wg, Ug = s.download()
    stats = s.stats()
    s.close()
    return st, (t_o, taus_o, speeds_o), (Wg, Ug, stats)


@pytest.mark.parametrize("schedule", [capi.SCHED_TILED, capi.SCHED_TWO_KERNEL])
def test_step_shock_tube_200_steps_bitwise(schedule):
    """Config 0: relativistic shock tube, 120k jittered/permuted tets, CFL 0.4, 200 steps,
    under both kernel schedules."""
    m = M.kuhn_box(200, 10, 10, length=(1.0, 0.05, 0.05), seed=SEED)
    assert (m.n_cells, m.n_faces) == (120000, 248200)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    st, (t_o, taus, speeds), (Wg, Ug, stats) = _run_both(m, W0, 0.4, 200, schedule=schedule)
    assert stats["steps"] == 200
    assert np.array_equal(st.U, Ug)
    assert np.array_equal(st.W, Wg)
    assert stats["max_speed"] == speeds[-1]
    assert stats["t"] == t_o
    for k in range(5):
        ref = np.abs(st.U[:, k]).sum()
        assert np.abs(st.U[:, k] - Ug[:, k]).sum() <= 1e-12 * ref


def test_step_smallest_meshes_bitwise():
    """One and two hexes (6 and 12 tets, 18 and 34 faces): every cell touches the
    boundary and one partial workgroup holds the whole mesh.  FP64: random states
    with jumps of up to seven decades, 30 steps, eager and graph, bitwise.  Mixed:
    states within a decade of each other, within 1e-5 L1 of the oracle."""
    rng = np.random.default_rng(SEED)
    for shape in ((1, 1, 1), (2, 1, 1)):
        m = M.kuhn_box(*shape, seed=SEED)
        W0 = random_prims(rng, m.n_cells, vmax=0.9)
        for graph in (False, True):
            st, (t_o, _, speeds), (Wg, Ug, stats) = _run_both(m, W0, 0.4, 30, graph=graph)
            assert np.array_equal(st.U, Ug) and np.array_equal(st.W, Wg), (shape, graph)
            assert stats["t"] == t_o and stats["max_speed"] == speeds[-1]
        W0 = random_prims(rng, m.n_cells, vmax=0.9, rho=(-0.5, 0.5), p=(-0.5, 0.5))
        st = O.State(m, W0, 5.0 / 3.0)
        st.run(0.4, m.h_min(), 30)
        s = capi.Solver(m, gamma=5.0 / 3.0, cfl=0.4, h_min=m.h_min(), precision=capi.PREC_MIXED,
                        guess=capi.GUESS_PREVIOUS)
        s.upload_W(W0)
        s.run(30)
        s.sync()
        Wm, Um = s.download()
        s.close()
        for k in range(5):
            assert np.abs(Um[:, k] - st.U[:, k]).sum() <= 1e-5 * np.abs(st.U[:, k]).sum(), (shape, k)


def test_step_periodic_shear_bitwise_eager_and_graph():
    """Periodic shear layer (W ~ 10, gamma 4/3) on 12^3 hexes, eager and graph replay agree bitwise."""
    m = M.kuhn_box(12, 12, 12, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
    W0 = M.init_state(m, M.INIT_SHEAR)
    st, (t_o, _, _), (Wg, Ug, _) = _run_both(m, W0, 0.4, 40, gamma=4.0 / 3.0, graph=False, reorder=False)
    assert np.array_equal(st.U, Ug) and np.array_equal(st.W, Wg)
    _, _, (Wg2, Ug2, _) = _run_both(m, W0, 0.4, 40, gamma=4.0 / 3.0, graph=True, reorder=True)
    assert np.array_equal(Ug, Ug2) and np.array_equal(Wg, Wg2)


def test_step_matches_host_tau_interface():
    """rhllc_step(tau) (the Hllc3dStab drop-in) returns the oracle's max signal speed."""
    m = M.kuhn_box(16, 4, 4, length=(1.0, 0.25, 0.25), seed=SEED)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    h = m.h_min()
    st = O.State(m, W0)
    st.run(0.4, h, 0)
    s = capi.Solver(m, cfl=0.4, h_min=h)
    s.upload_W(W0)
    tau = 0.4 * h
    for _ in range(10):
        sp_o = st.step(tau)
        sp_g = s.step(tau)
        assert sp_o == sp_g
        tau = 0.4 * h / sp_o
    _, Ug = s.download()
    assert np.array_equal(st.U, Ug)
    s.close()


def test_mixed_precision_blast_l1():
    """Mixed FP32-flux / FP64-c2p against the FP64 step on a small blast wave."""
    m = M.kuhn_box(24, 24, 24, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    h = m.h_min()
    res = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = capi.Solver(m, cfl=0.4, h_min=h, precision=prec, guess=guess)
        s.upload_W(W0)
        s.run(50)
        s.sync()
        res[prec] = s.download()[1]
        s.close()
    U64, U32 = res[capi.PREC_FP64], res[capi.PREC_MIXED]
    V = m.cells["V"]
    l1 = [float((np.abs(U32[:, k] - U64[:, k]) * V).sum() / (np.abs(U64[:, k]) * V).sum()) for k in range(5)]
    print("mixed vs fp64 rel L1 per conserved field:", l1)
    assert max(l1) < 1e-5


@pytest.mark.parametrize("schedule", [capi.SCHED_TILED, capi.SCHED_TWO_KERNEL])
def test_step_radiation_source_bitwise(schedule):
    """Full-form coupling: hydro step, U += tau*G(W) (M1 closure), c2p -- as radRHD_main.cpp:99-132."""
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    Ur, Fr, ka, ks = M.init_radiation(m.n_cells, seed=SEED)
    h = m.h_min()
    st = O.State(m, W0)
    st.run_rad(0.4, h, 30, Ur, Fr, ka, ks)
    s = capi.Solver(m, cfl=0.4, h_min=h, schedule=schedule)
    s.upload_W(W0)
    s.set_radiation(Ur, Fr, ka, ks)
    s.run(30)
    s.sync()
    W, U = s.download()
    s.close()
    assert np.array_equal(st.U, U) and np.array_equal(st.W, W)


# ------------------------------------------------ EOS TAUB (the reference's shipped EOS)
def test_eval_taub_bitwise_against_oracle_and_reference_vectors():
    import os

    T = capi.EOS_TAUB
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "taub.npz"), allow_pickle=False)
    Wc, U = capi.eval_p2c(d["p2c_W_in"], eos=T)
    assert np.array_equal(Wc, d["p2c_W_out"]) and np.array_equal(U, d["p2c_U"])
    W, codes, _ = capi.eval_c2p(d["c2p_U"], W_init=d["c2p_W_init"], eos=T)
    assert np.array_equal(codes, d["c2p_codes"]) and np.array_equal(W, d["c2p_W_out"])
    U2, W2, c2 = capi.eval_pressure_fix(d["fix_U_in"], d["fix_W_in"], eos=T)
    assert np.array_equal(c2, d["fix_codes"]) and np.array_equal(U2, d["fix_U_out"]) and np.array_equal(W2, d["fix_W_out"])
    F, sp = capi.eval_flux(d["flux_UL"], d["flux_UR"], d["flux_WL"], d["flux_WR"], d["flux_nS"], eos=T)
    assert np.array_equal(F, d["flux_F"]) and np.array_equal(sp, d["flux_speed"])
    # and against the oracle on fresh states
    rng = np.random.default_rng(SEED + 7)
    n = 20000
    WL, WR = random_prims(rng, n), random_prims(rng, n)
    _, UL = O.p2c(WL, eos=T)
    _, UR = O.p2c(WR, eos=T)
    nS = random_faces(rng, n)
    Fo, so = O.flux(UL, UR, WL, WR, nS, eos=T)
    Fg, sg = capi.eval_flux(UL, UR, WL, WR, nS, eos=T)
    assert np.array_equal(Fo, Fg) and np.array_equal(so, sg)
    Wo, co, io = O.c2p(UL, eos=T)
    Wg, cg, ig = capi.eval_c2p(UL, eos=T)
    assert np.array_equal(co, cg) and np.array_equal(io, ig) and np.array_equal(Wo, Wg)


def test_step_taub_bitwise():
    """Shock tube and blast with the Taub-Mathews EOS: GPU FP64 step bitwise the oracle
    (itself bitwise the reference built with EOS TAUB), and the reference's own
    200-step trajectory from tests/golden/taub.npz."""
    import os

    T = capi.EOS_TAUB
    m = M.kuhn_box(40, 4, 4, length=(1.0, 0.1, 0.1), seed=SEED)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    st, (t_o, taus, speeds), (Wg, Ug, stats) = _run_both(m, W0, 0.4, 150, eos=T)
    assert np.array_equal(st.U, Ug) and np.array_equal(st.W, Wg)
    assert stats["max_speed"] == speeds[-1] and stats["t"] == t_o
    m = M.kuhn_box(10, 10, 10, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST, params=[0.5, 0.5, 0.5, 0.25, 1.0, 1000.0, 1.0])
    st, _, (Wg, Ug, _) = _run_both(m, W0, 0.4, 60, eos=T)
    assert np.array_equal(st.U, Ug) and np.array_equal(st.W, Wg)
    d = np.load(os.path.join(os.path.dirname(__file__), "golden", "taub.npz"), allow_pickle=False)
    m = M.Mesh(d["traj_geo_cells"].view(M.GEO_CELL_DTYPE).copy(), d["traj_geo_faces"].view(M.GEO_FACE_DTYPE).copy())
    s = capi.Solver(m, cfl=float(d["traj_cfl"]), h_min=float(d["traj_h_min"]), eos=T)
    s.upload_W(d["traj_W0"])
    s.run(200)
    s.sync()
    Wg, Ug = s.download()
    s.close()
    assert np.array_equal(Ug, d["traj_U"]) and np.array_equal(Wg, d["traj_W"])


def test_mixed_taub_blast_l1():
    T = capi.EOS_TAUB
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    out = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = capi.Solver(m, cfl=0.4, precision=prec, guess=guess, eos=T)
        s.upload_W(W0)
        s.run(60)
        s.sync()
        out[prec] = (s.download()[1], s.stats())
        s.close()
    U64, U32 = out[capi.PREC_FP64][0], out[capi.PREC_MIXED][0]
    assert out[capi.PREC_MIXED][1]["fix_fail"] == 0 and sum(out[capi.PREC_MIXED][1]["c2p_fail"]) == 0
    for k in range(5):
        ref = np.abs(U64[:, k]).sum()
        assert np.abs(U32[:, k] - U64[:, k]).sum() <= 1e-5 * ref, k
    with pytest.raises(capi.RhllcError):
        capi.Solver(m, precision=capi.PREC_MIXED, guess=capi.GUESS_MLP, eos=T)


def test_mixed_radiation_l1():
    """Mixed step with the radiation source (the compact M1 form of the mixed path)
    against the FP64 step with the reference's tensor form: per-field L1 <= 1e-5."""
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    W0 = M.init_state(m, M.INIT_BLAST)
    Ur, Fr, ka, ks = M.init_radiation(m.n_cells, seed=SEED)
    Fr[:50] = 0.0  # the vanishing-flux branch
    Fr[50:100] *= 3.0 / np.maximum(np.linalg.norm(Fr[50:100], axis=1, keepdims=True), 1e-30) * Ur[50:100, None]  # |F| > U
    out = {}
    for prec, guess in ((capi.PREC_FP64, capi.GUESS_REFERENCE), (capi.PREC_MIXED, capi.GUESS_PREVIOUS)):
        s = capi.Solver(m, cfl=0.4, precision=prec, guess=guess)
        s.upload_W(W0)
        s.set_radiation(Ur, Fr, ka, ks)
        s.run(40)
        s.sync()
        out[prec] = (s.download()[1], s.stats())
        s.close()
    U64, U32 = out[capi.PREC_FP64][0], out[capi.PREC_MIXED][0]
    assert out[capi.PREC_MIXED][1]["fix_fail"] == 0
    for k in range(5):
        ref = np.abs(U64[:, k]).sum()
        assert np.abs(U32[:, k] - U64[:, k]).sum() <= 1e-5 * ref, k
