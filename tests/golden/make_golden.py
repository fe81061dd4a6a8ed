#!/usr/bin/env python3
"""Writes the golden fixtures of the RHLLC step from the reference's own code.

Run in a container where the reference tree exists and `make -C oracle ref`
has built oracle/_ref/libref_rhllc.so (the reference's hot-path sources,
EOS IDEAL, gamma 5/3, strict IEEE).  Inputs are seeded; outputs are what the
compiled reference returns.  The fixtures are data only (inputs + outputs):

  c2p.npz        GetPhysValueStab        U -> W, return code
  p2c.npz        GetConvValueStab        W -> (W clamped, U)
  fix.npz        PhysPressureFix         (U, W) -> (U with E rewritten, W, code)
  rotation.npz   GetRotationMatrix       n -> T
  flux.npz       GetFluxStab             (UL, UR, WL, WR, n, S) -> (F, max(lL, lR)), all 4 branches
  traj_tiny.npz  Hllc3dStab loop         8x2x2-hex jittered/permuted mesh, 200 steps
  shock_tube_200.json  config 0 (200x10x10 hexes, 120k tets) after 200 steps: per-field sums and hashes
  brick.vtk.xz, sphere.vtk.xz  the reference's test meshes (tests/data), compressed
  brick_cone.json  the reference's RHLLC test on brick.vtk to T = 0.5: geometry and state hashes
  taub.npz, brick_cone_taub.json  the pointwise functions, a tiny trajectory and the brick test
                 from the reference built with EOS TAUB (its shipped setting)
  rcf_H.dat.xz, table.npz  the reference's cooling-function table (tab_func), compressed, and
                 what its ReadTableFunc and TableFunc::operator() return on it

``python make_golden.py --vtk`` rewrites only the brick/sphere files,
``--taub`` only the EOS TAUB files, ``--table`` only the table files.
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from radrhd_solver_amd import mesh as M  # noqa: E402

SEED = 20261202


def prims(rng, n, rho=(-4, 2), pr=(-6, 2), wmax=50.0):
    """Config-3 sampling box: log-uniform rho, p/rho, Lorentz factor W in [1, wmax], isotropic v."""
    r = 10 ** rng.uniform(*rho, n)
    p = r * 10 ** rng.uniform(*pr, n)
    W = np.exp(rng.uniform(0, np.log(wmax), n))
    v = np.sqrt(1 - 1 / W**2)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    return np.column_stack([r, d * v[:, None], p])


def unit(rng, n):
    x = rng.normal(size=(n, 3))
    return x / np.linalg.norm(x, axis=1)[:, None]


REF_ROOT = "/root/reference"
REF_DATA = os.path.join(REF_ROOT, "tests", "data")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def vtk_fixtures():
    """The reference's own test meshes (tests/data/brick.vtk, sphere.vtk), xz-compressed,
    and the reference's RHLLC test on brick.vtk (tests/src/off_rhllc_test.cpp:22-45):
    GEOMETRY_TYPE Cone boundaries and initial state, CFL 0.7, h_min 1e-5, run to T = 0.5.

    The geometry arrays come from the oracle's restatement of the VTK-based
    preprocessing (oracle/mesh_geo.py; VTK itself is not available), the
    geometry records from the reference's own ConvertBinToGeo, the run from
    the reference's own Hllc3dStab driven like rhllc_main.cpp:34-56."""
    import lzma

    from oracle import mesh_geo as G

    for name in ("brick", "sphere"):
        with open(os.path.join(REF_DATA, name + ".vtk"), "rb") as f:
            raw = f.read()
        with open(os.path.join(HERE, name + ".vtk.xz"), "wb") as f:
            f.write(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME))

    P, T = G.read_legacy_vtk(os.path.join(REF_DATA, "brick.vtk"))
    nb, nrm, ar, vol, cen = G.build_data_from_vtk(P, T, "Cone")
    cells, faces = O.ref_convert_bin_to_geo(nb, nrm, ar, vol, cen)
    # Writable views of the raw bytes: a structured copy would not keep the padding bytes.
    m = M.Mesh(np.frombuffer(bytearray(cells.tobytes()), M.GEO_CELL_DTYPE),
               np.frombuffer(bytearray(faces.tobytes()), M.GEO_FACE_DTYPE))
    W_init, W_src, U_src = O.ref_cone_states()
    W0 = np.tile(W_init, (m.n_cells, 1))
    cfl, h_min, t_end = 0.7, 1e-5, 0.5
    _, _, taus, _ = O.ref_run(m, W0, cfl, h_min, 2000)
    # rhllc_main.cpp:34-56: while (t < T) { step(tau); t += tau; tau = GetTimeStep }
    t, steps = 0.0, 0
    while t < t_end:
        t += taus[steps]
        steps += 1
    Wt, Ut, taus, speeds = O.ref_run(m, W0, cfl, h_min, steps)
    t = 0.0
    for tau in taus:  # the driver's running sum, in step order
        t += float(tau)
    rec = {
        "source": "reference tests/src/off_rhllc_test.cpp:22-45 on tests/data/brick.vtk",
        "n_points": int(P.shape[0]), "n_cells": m.n_cells, "n_faces": m.n_faces,
        "neighbors_sha256": sha(nb), "normals_sha256": sha(nrm), "areas_sha256": sha(ar),
        "volumes_sha256": sha(vol), "centers_sha256": sha(cen),
        "geo_cells_sha256": sha(m.cells), "geo_faces_sha256": sha(m.faces),
        "bound_counts": {str(c): int((nb == c).sum()) for c in (-2, -3, -5)},
        "W_init": W_init.tolist(), "W_source": W_src.tolist(), "U_source": U_src.tolist(),
        "cfl": cfl, "h_min": h_min, "T": t_end, "steps": steps, "t": t,
        "taus_sha256": sha(taus), "last_speed": float(speeds[-1]),
        "U_abs_sum": [float(np.abs(Ut[:, j]).sum()) for j in range(5)],
        "W_abs_sum": [float(np.abs(Wt[:, j]).sum()) for j in range(5)],
        "U_sha256": sha(Ut), "W_sha256": sha(Wt),
    }
    with open(os.path.join(HERE, "brick_cone.json"), "w") as f:
        json.dump(rec, f, indent=1)


def taub_fixtures():
    """The same pointwise functions and a tiny trajectory from the reference
    compiled with EOS TAUB (its shipped include/hd_config.h:20), and the
    brick.vtk test of vtk_fixtures() with EOS TAUB."""
    T = O.EOS_TAUB
    rng = np.random.default_rng(SEED + 100)
    W = prims(rng, 500)
    W[:25, 1:4] *= 1.0 / np.linalg.norm(W[:25, 1:4], axis=1)[:, None] * 1.0001
    Wc, U = O.ref_p2c(W, eos=T)
    out = {"p2c_W_in": W, "p2c_W_out": Wc, "p2c_U": U}
    W = prims(rng, 2000)
    _, U = O.ref_p2c(W, eos=T)
    U[:100, 4] = np.sqrt(U[:100, 0] ** 2 + (U[:100, 1:4] ** 2).sum(1)) * (1 - 1e-3)  # Q < 0: code 1
    U[100:200, 4] *= 0.999999  # near-admissible states
    W_init = np.full((U.shape[0], 5), -7.0)
    Wo, codes = O.ref_c2p(U, W_init, eos=T)
    out.update(c2p_U=U, c2p_W_init=W_init, c2p_W_out=Wo, c2p_codes=codes)
    bad = codes != 0
    Uf, Wf, cf = O.ref_pressure_fix(U[bad], Wo[bad], eos=T)
    out.update(fix_U_in=U[bad], fix_W_in=Wo[bad], fix_U_out=Uf, fix_W_out=Wf, fix_codes=cf)
    k = 3000
    WL = prims(rng, k, rho=(-2, 2), pr=(-3, 2), wmax=10)
    WR = prims(rng, k, rho=(-2, 2), pr=(-3, 2), wmax=10)
    WL[:300, 1:4] = [0.99, 0, 0]
    WR[:300, 1:4] = [0.99, 0, 0]
    WL[300:600, 1:4] = [-0.99, 0, 0]
    WR[300:600, 1:4] = [-0.99, 0, 0]
    _, UL = O.ref_p2c(WL, eos=T)
    _, UR = O.ref_p2c(WR, eos=T)
    nn = unit(rng, k)
    nn[600:606] = np.eye(3).tolist() + (-np.eye(3)).tolist()
    nS = np.column_stack([nn, rng.uniform(0.01, 2.0, k)])
    F, sp = O.ref_flux(UL, UR, WL, WR, nS, eos=T)
    out.update(flux_UL=UL, flux_UR=UR, flux_WL=WL, flux_WR=WR, flux_nS=nS, flux_F=F, flux_speed=sp)
    m = M.kuhn_box(8, 2, 2, length=(1.0, 0.25, 0.25), seed=SEED)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    h = m.h_min()
    Wt, Ut, taus, speeds = O.ref_run(m, W0, 0.4, h, 200, eos=T)
    out.update(traj_geo_cells=m.cells.view(np.uint8).reshape(-1), traj_geo_faces=m.faces.view(np.uint8).reshape(-1),
               traj_W0=W0, traj_cfl=0.4, traj_h_min=h, traj_W=Wt, traj_U=Ut, traj_taus=taus, traj_speeds=speeds)
    np.savez(os.path.join(HERE, "taub.npz"), **out)

    # brick.vtk with EOS TAUB, the configuration the reference ships
    with open(os.path.join(HERE, "brick_cone.json")) as f:
        ideal = json.load(f)
    from oracle import mesh_geo as G

    P, Tt = G.read_legacy_vtk(os.path.join(REF_DATA, "brick.vtk"))
    nb, nrm, ar, vol, cen = G.build_data_from_vtk(P, Tt, "Cone")
    cells, faces = O.ref_convert_bin_to_geo(nb, nrm, ar, vol, cen)
    m = M.Mesh(np.frombuffer(bytearray(cells.tobytes()), M.GEO_CELL_DTYPE),
               np.frombuffer(bytearray(faces.tobytes()), M.GEO_FACE_DTYPE))
    W0 = np.tile(np.array(ideal["W_init"]), (m.n_cells, 1))
    _, _, taus, _ = O.ref_run(m, W0, ideal["cfl"], ideal["h_min"], 2000, eos=T)
    t, steps = 0.0, 0
    while t < ideal["T"]:
        t += taus[steps]
        steps += 1
    Wt, Ut, taus, speeds = O.ref_run(m, W0, ideal["cfl"], ideal["h_min"], steps, eos=T)
    t = 0.0
    for tau in taus:
        t += float(tau)
    rec = {"source": ideal["source"] + ", reference built with EOS TAUB", "steps": steps, "t": t,
           "taus_sha256": sha(taus), "last_speed": float(speeds[-1]), "U_sha256": sha(Ut), "W_sha256": sha(Wt),
           "U_abs_sum": [float(np.abs(Ut[:, j]).sum()) for j in range(5)]}
    with open(os.path.join(HERE, "brick_cone_taub.json"), "w") as f:
        json.dump(rec, f, indent=1)


def table_fixtures():
    """The reference's cooling-function table (tab_func/rcf_H.dat), xz-compressed, and
    table.npz: what the reference's own ReadTableFunc (lib/files_sys/src/reader_txt.cpp:59-118)
    reads from it (header, data hash) and what its TableFunc::operator()
    (lib/grids/src/solvers_struct.cpp:10-58) returns at seeded points: inside the table,
    on and between its nodes, below both minima, above max_x, and above max_y (the
    extrapolation branch)."""
    import lzma

    src = os.path.join(REF_ROOT, "tab_func", "rcf_H.dat")
    with open(src, "rb") as f:
        raw = f.read()
    with open(os.path.join(HERE, "rcf_H.dat.xz"), "wb") as f:
        f.write(lzma.compress(raw, preset=9 | lzma.PRESET_EXTREME))
    t = O.ref_table_read(src)
    rng = np.random.default_rng(SEED + 31)
    mnx, mxx, sx, mny, mxy, sy = t.hdr
    n = 4000
    x = rng.uniform(mnx - 3, mxx + 3, n)
    y = rng.uniform(mny - 2, mxy + 3, n)
    k = np.arange(200)
    x[:200] = mnx + sx * (k % t.data.shape[0])  # node values of the axes
    y[:200] = mny + sy * (k % t.data.shape[1])
    x[200:300] = np.round(rng.uniform(mnx, mxx, 100)) + 0.5  # half-integers: round() ties
    y[200:300] = np.round(rng.uniform(mny, mxy, 100)) + 0.5
    y[300:400] = mxy + rng.uniform(0, 1e-3, 100)  # just above max_y
    y[400:420] = mxy  # exactly max_y
    x[420:440], y[420:440] = mnx, mny
    out = O.ref_table_lookup(t, x, y)
    np.savez(os.path.join(HERE, "table.npz"), hdr=t.hdr, shape=np.array(t.data.shape), data_sha256=sha(t.data),
             x=x, y=y, value=out)


def main():
    if not O.have_ref():
        sys.exit("oracle/_ref/libref_rhllc.so is not built (make -C oracle ref)")
    if sys.argv[1:] == ["--vtk"]:
        vtk_fixtures()
        print("wrote VTK fixtures to", HERE)
        return
    if sys.argv[1:] == ["--table"]:
        table_fixtures()
        print("wrote table fixtures to", HERE)
        return
    if sys.argv[1:] == ["--taub"]:
        taub_fixtures()
        print("wrote EOS TAUB fixtures to", HERE)
        return
    rng = np.random.default_rng(SEED)
    out = {}

    # p2c, with some |v| >= 1 inputs (clamped to 0.9999 by the reference)
    W = prims(rng, 500)
    W[:25, 1:4] *= 1.0 / np.linalg.norm(W[:25, 1:4], axis=1)[:, None] * 1.0001
    Wc, U = O.ref_p2c(W)
    np.savez(os.path.join(HERE, "p2c.npz"), W_in=W, W_out=Wc, U=U)

    # c2p: valid states plus states with E lowered below sqrt(m^2 + D^2) (code 1)
    W = prims(rng, 1500)
    _, U = O.ref_p2c(W)
    U[:100, 4] = np.sqrt(U[:100, 0] ** 2 + (U[:100, 1:4] ** 2).sum(1)) * (1 - 1e-6)
    W0 = np.full_like(U, -3.0)
    Wr, codes = O.ref_c2p(U, W0)
    np.savez(os.path.join(HERE, "c2p.npz"), U=U, W_init=W0, W_out=Wr, codes=codes)

    # PhysPressureFix on the failing states and on low-pressure states
    Wf = prims(rng, 100, pr=(-9, -6))
    _, Uf = O.ref_p2c(Wf)
    Uf[:50] = U[:50]
    Wf_in = np.full_like(Uf, -5.0)
    Uo, Wo, fc = O.ref_pressure_fix(Uf, Wf_in)
    np.savez(os.path.join(HERE, "fix.npz"), U_in=Uf, W_in=Wf_in, U_out=Uo, W_out=Wo, codes=fc)

    # rotations, including n = +-z exactly and near it
    n = unit(rng, 60)
    n = np.vstack([n, [[0, 0, 1], [0, 0, -1], [0, 0, 1 - 1e-13], [1, 0, 0]]])
    n[-2] /= np.linalg.norm(n[-2])
    np.savez(os.path.join(HERE, "rotation.npz"), n=n, T=O.ref_rotation(n))

    # fluxes: random pairs (subsonic HLLC, both us signs) + supersonic pairs
    k = 1500
    WL = prims(rng, k, rho=(-2, 2), pr=(-3, 2), wmax=10)
    WR = prims(rng, k, rho=(-2, 2), pr=(-3, 2), wmax=10)
    nn = unit(rng, k)
    nn[:6] = [[0, 0, 1], [0, 0, -1], [1, 0, 0], [0, 1, 0], [-1, 0, 0], [0, -1, 0]]
    cold = slice(6, 206)  # cold supersonic streams along +n and -n
    WL[cold, 4] = WL[cold, 0] * 1e-4
    WR[cold, 4] = WR[cold, 0] * 1e-4
    WL[6:106, 1:4] = 0.95 * nn[6:106]
    WR[6:106, 1:4] = 0.95 * nn[6:106]
    WL[106:206, 1:4] = -0.95 * nn[106:206]
    WR[106:206, 1:4] = -0.95 * nn[106:206]
    WR[206:256] = WL[206:256]  # identical states
    _, UL = O.ref_p2c(WL)
    _, UR = O.ref_p2c(WR)
    nS = np.column_stack([nn, rng.uniform(0.01, 2.0, k)])
    F, sp = O.ref_flux(UL, UR, WL, WR, nS)
    np.savez(os.path.join(HERE, "flux.npz"), UL=UL, UR=UR, WL=WL, WR=WR, nS=nS, F=F, speed=sp)

    # tiny trajectory: 8x2x2 hexes, shock tube, 200 steps
    m = M.kuhn_box(8, 2, 2, length=(1.0, 0.25, 0.25), seed=SEED)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    h = m.h_min()
    Wt, Ut, taus, speeds = O.ref_run(m, W0, 0.4, h, 200)
    np.savez(
        os.path.join(HERE, "traj_tiny.npz"),
        geo_cells=m.cells.view(np.uint8).reshape(-1), geo_faces=m.faces.view(np.uint8).reshape(-1),
        W0=W0, cfl=0.4, h_min=h, W=Wt, U=Ut, taus=taus, speeds=speeds,
    )

    # config 0 after 200 steps (the fixture is too large to store; sums and hashes are)
    m = M.kuhn_box(200, 10, 10, length=(1.0, 0.05, 0.05), seed=SEED)
    W0 = M.init_state(m, M.INIT_SHOCK_TUBE)
    h = m.h_min()
    Wt, Ut, taus, speeds = O.ref_run(m, W0, 0.4, h, 200)
    rec = {
        "mesh": {"hexes": [200, 10, 10], "length": [1.0, 0.05, 0.05], "seed": SEED, "n_cells": m.n_cells,
                 "n_faces": m.n_faces, "geo_cells_sha256": hashlib.sha256(m.cells.tobytes()).hexdigest(),
                 "geo_faces_sha256": hashlib.sha256(m.faces.tobytes()).hexdigest()},
        "h_min": h, "cfl": 0.4, "steps": 200,
        "t": float(taus.sum()), "last_speed": float(speeds[-1]), "last_tau": float(taus[-1]),
        "U_abs_sum": [float(np.abs(Ut[:, j]).sum()) for j in range(5)],
        "W_abs_sum": [float(np.abs(Wt[:, j]).sum()) for j in range(5)],
        "U_sha256": hashlib.sha256(np.ascontiguousarray(Ut).tobytes()).hexdigest(),
        "W_sha256": hashlib.sha256(np.ascontiguousarray(Wt).tobytes()).hexdigest(),
    }
    with open(os.path.join(HERE, "shock_tube_200.json"), "w") as f:
        json.dump(rec, f, indent=1)
    vtk_fixtures()
    taub_fixtures()
    table_fixtures()
    print("wrote fixtures to", HERE)


if __name__ == "__main__":
    main()
