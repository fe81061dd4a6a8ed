"""Synthetic Kuhn-tet meshes in the reference's record format.

Checks the ConvertBinToGeo numbering/orientation rule
(lib/geometry/src/convert_bin_to_geo.cpp:3-48), the geometry, periodic pairing,
h_min (utils/src/get_average_volume.cpp:31-46) and the binary file formats
(lib/files_sys/include/reader_bin.h:113-135).
"""
import os

import numpy as np
import pytest

from radrhd_solver_amd import mesh as M

SEED = 20261202


def decode(m):
    return M.face_id_l(m.faces), M.face_id_r(m.faces)


@pytest.mark.parametrize("dims,periodic", [((5, 3, 2), 0), ((4, 4, 4), 7), ((6, 2, 3), 1), ((3, 5, 4), 6)])
def test_counts_and_convert_bin_to_geo_rule(dims, periodic):
    m = M.kuhn_box(*dims, periodic=periodic, seed=SEED)
    nx, ny, nz = dims
    N = 6 * nx * ny * nz
    B = sum(4 * a * b for ax, (a, b) in enumerate(((ny, nz), (nx, nz), (nx, ny))) if not (periodic >> ax) & 1)
    assert m.n_cells == N and m.n_faces == (4 * N + B) // 2
    idl, idr = decode(m)
    ids = m.cells["id_faces"]
    sign = m.cells["sign_n"]
    # every slot points at a face that points back, with the sign bit telling the side
    for j in range(4):
        f = ids[:, j]
        left = (sign >> j) & 1
        assert np.array_equal(np.where(left == 1, idl[f], idr[f]), np.arange(N))
    # faces are numbered in order of their first visit (cell-major, slot order)
    _, first = np.unique(ids.reshape(-1), return_index=True)
    assert (np.diff(first) > 0).all()
    # so id_l is the smaller cell index of an interior face
    inner = idr >= 0
    assert (idl[inner] < idr[inner]).all()
    assert (idr[~inner] == -1).all()
    # is_bound counts the cell's boundary faces
    nbound = np.zeros(N, np.int64)
    np.add.at(nbound, idl[~inner], 1)
    assert np.array_equal(M.cell_is_bound(m.cells), nbound & 3)


def test_orientation_volume_and_closure():
    m = M.kuhn_box(7, 5, 4, length=(1.4, 1.0, 0.8), seed=SEED)
    idl, idr = decode(m)
    c = m.cells["center"]
    n = m.faces["n"]
    inner = idr >= 0
    # normals point from id_l towards id_r; boundary normals point out of the box
    assert (np.einsum("ij,ij->i", n[inner], c[idr[inner]] - c[idl[inner]]) > 0).all()
    out = c[idl[~inner]] - np.array([0.7, 0.5, 0.4])
    assert (np.einsum("ij,ij->i", n[~inner], out) > 0).all()
    assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-14)
    assert np.isclose(m.cells["V"].sum(), 1.4 * 1.0 * 0.8, rtol=1e-13)
    acc = np.zeros((m.n_cells, 3))
    for j in range(4):
        f = m.cells["id_faces"][:, j]
        s = np.where((m.cells["sign_n"] >> j) & 1, 1.0, -1.0)
        acc += s[:, None] * n[f] * m.faces["S"][f][:, None]
    assert np.abs(acc).max() < 1e-14


def test_jitter_permutation_are_seeded():
    a = M.kuhn_box(4, 3, 2, seed=1)
    b = M.kuhn_box(4, 3, 2, seed=1)
    c = M.kuhn_box(4, 3, 2, seed=2)
    assert a.cells.tobytes() == b.cells.tobytes() and a.faces.tobytes() == b.faces.tobytes()
    assert a.cells.tobytes() != c.cells.tobytes()
    u = M.kuhn_box(4, 3, 2, jitter=0.0, permute=False)
    # unjittered, unpermuted: every tet has volume h^3/6
    assert np.allclose(u.cells["V"], (1 / 4) * (1 / 3) * (1 / 2) / 6, rtol=1e-13)


def test_periodic_partners_have_equal_area():
    m = M.kuhn_box(4, 4, 4, periodic=7, seed=SEED)
    idl, idr = decode(m)
    assert (idr >= 0).all()
    # every cell has 4 distinct neighbours on a fully periodic lattice
    nb = np.where((m.cells["sign_n"][:, None] >> np.arange(4)) & 1, idr[m.cells["id_faces"]], idl[m.cells["id_faces"]])
    assert (nb != np.arange(m.n_cells)[:, None]).all()


def test_h_min_formula():
    m = M.kuhn_box(5, 4, 3, seed=SEED)
    S = m.faces["S"][m.cells["id_faces"]]
    surf = ((S[:, 0] + S[:, 1]) + S[:, 2]) + S[:, 3]
    assert m.h_min() == float(np.min(0.75 * m.cells["V"] / surf))


def test_record_files_round_trip(tmp_path):
    m = M.kuhn_box(3, 2, 2, seed=SEED)
    fp, cp = str(tmp_path / "geo_faces.bin"), str(tmp_path / "geo_cells.bin")
    m.write(fp, cp)
    assert os.path.getsize(fp) == 4 + 40 * m.n_faces and os.path.getsize(cp) == 4 + 56 * m.n_cells
    r = M.Mesh.read(fp, cp)
    assert r.cells.tobytes() == m.cells.tobytes() and r.faces.tobytes() == m.faces.tobytes()


def test_initial_states():
    m = M.kuhn_box(20, 2, 2, length=(1.0, 0.1, 0.1), seed=SEED)
    W = M.init_state(m, M.INIT_SHOCK_TUBE)
    left = m.cells["center"][:, 0] < 0.5
    assert (W[left, 0] == 10.0).all() and (W[left, 4] == 13.33).all()
    assert (W[~left, 0] == 1.0).all() and (W[~left, 4] == 1e-6).all()
    mb = M.kuhn_box(8, 8, 8, seed=SEED)
    Wb = M.init_state(mb, M.INIT_BLAST)
    r = np.linalg.norm(mb.cells["center"] - 0.5, axis=1)
    assert np.array_equal(Wb[:, 4], np.where(r < 0.1, 1000.0, 1.0))
    ms = M.kuhn_box(8, 8, 8, origin=(-0.5, -0.5, -0.5), periodic=7, seed=SEED)
    Ws = M.init_state(ms, M.INIT_SHEAR)
    y = ms.cells["center"][:, 1]
    assert np.allclose(Ws[:, 1], 0.995 * np.tanh(y / 0.01))
    assert np.abs(Ws[:, 2]).max() <= 4e-3 + 1e-4
    Ur, Fr, ka, ks = M.init_radiation(1000, seed=SEED)
    assert (Ur >= 1e-3 * 0.999).all() and (Ur <= 1.0 * 1.001).all()
    assert (np.linalg.norm(Fr, axis=1) <= Ur * (1 + 1e-6)).all()
    assert ka.min() >= 0.1 and ks.max() <= 10.0


def test_curve_order_is_a_permutation_with_faces_grouped():
    m = M.kuhn_box(6, 5, 4, seed=SEED)
    cp, fp = M.curve_order(m)
    assert np.array_equal(np.sort(cp), np.arange(m.n_cells))
    assert np.array_equal(np.sort(fp), np.arange(m.n_faces))
    new_of_old = np.empty_like(cp)
    new_of_old[cp] = np.arange(m.n_cells)
    idl, idr = M.face_id_l(m.faces), M.face_id_r(m.faces)
    owner = new_of_old[idl]
    inner = idr >= 0
    owner[inner] = np.minimum(owner[inner], new_of_old[idr[inner]])
    assert (np.diff(owner[fp]) >= 0).all()
    # ties keep the old face order
    same = np.diff(owner[fp]) == 0
    assert (np.diff(fp)[same] > 0).all()


def test_curve_order_walks_between_neighbours():
    # On a 16^3 box (24,576 tets of 1/16 edge) consecutive cells of a Hilbert
    # order are neighbours or near it; a Morton order jumps by up to the box.
    m = M.kuhn_box(16, 16, 16, seed=SEED)
    cp, _ = M.curve_order(m)
    x = m.cells["center"][cp]
    step = np.linalg.norm(np.diff(x, axis=0), axis=1) * 16
    assert step.max() < 2.0 and step.mean() < 0.6


def test_generator_rejects_bad_boxes():
    with pytest.raises(ValueError):
        M.kuhn_box(0, 2, 2)
    with pytest.raises(ValueError):
        M.kuhn_box(1, 2, 2, periodic=1)
    with pytest.raises(ValueError):
        M.kuhn_box(200, 200, 200)  # 48M cells > 2^25 (id_l is a 25-bit field)
