"""The reference's mesh preprocessing without VTK: legacy-VTK reader, geometry
arrays and geo records (CPU).

The reference builds its solver geometry from VTK files
(BuildDataFromVTK, lib/geometry/src/make_internal_format.cpp:272-336, then
ConvertBinToGeo, lib/geometry/src/convert_bin_to_geo.cpp:3-48).  Checks here:

* the oracle restatement (oracle/mesh_geo.py) against what the mathematics
  fixes -- volumes that tile the box, closed cells, outward normals, the
  insphere centre, mutual neighbours with opposite normals, boundary codes by
  the plane a face lies on, the VTK face table's orientation, the orientation
  polynomial equal to the triple product, and a two-cell mesh by hand;
* the mesh library (product) bitwise against the oracle on the reference's own
  test meshes (tests/data/brick.vtk, sphere.vtk, committed compressed), and
  its ConvertBinToGeo byte for byte against the reference's own, compiled
  (oracle/_ref) where available and always against the committed hashes;
* the reader on every legacy encoding it accepts, and its errors;
* the reference's own RHLLC test on brick.vtk (tests/src/off_rhllc_test.cpp)
  through the oracle step to T = 0.5, against the compiled reference's hashes.
"""
import hashlib
import json
import lzma
import os
import struct

import numpy as np
import pytest

import oracle as O
from oracle import mesh_geo as G
from radrhd_solver_amd import mesh as M

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")


def sha(a) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.fixture(scope="module")
def vtk_files(tmp_path_factory):
    d = tmp_path_factory.mktemp("vtk")
    out = {}
    for name in ("brick", "sphere"):
        p = d / (name + ".vtk")
        with lzma.open(os.path.join(GOLD, name + ".vtk.xz")) as f:
            p.write_bytes(f.read())
        out[name] = str(p)
    return out


@pytest.fixture(scope="module")
def brick(vtk_files):
    P, T = G.read_legacy_vtk(vtk_files["brick"])
    return P, T, G.build_data_from_vtk(P, T, "Cone")


@pytest.fixture(scope="module")
def sphere(vtk_files):
    P, T = G.read_legacy_vtk(vtk_files["sphere"])
    return P, T, G.build_data_from_vtk(P, T, "Cone")


def gold_brick():
    with open(os.path.join(GOLD, "brick_cone.json")) as f:
        return json.load(f)


# ------------------------------------------------------------ oracle pins
def _pins(P, T, arrays):
    nb, nrm, ar, vol, cen = arrays
    n = T.shape[0]
    S = ar.reshape(n, 4)
    X = P[T]  # [n, 4, 3]
    assert (vol > 0).all()
    # closed cells: sum_j S_j n_j = 0
    closure = (S[:, :, None] * nrm).sum(1)
    assert np.abs(closure).max() <= 1e-12 * S.sum(1).max()
    # unit normals pointing away from the cell
    assert np.allclose(np.linalg.norm(nrm, axis=2), 1.0, atol=1e-14)
    for j, f in enumerate(G.VTK_TETRA_FACES):
        fc = X[:, list(f)].mean(1)
        assert ((nrm[:, j] * (fc - X.mean(1))).sum(1) > 0).all()
        # the centre is the insphere centre: every face plane at distance 3V / sum S
        dist = (nrm[:, j] * (X[:, f[0]] - cen)).sum(1)
        np.testing.assert_allclose(dist, 3 * vol / S.sum(1), rtol=1e-8)
        # the area is half the cross product of the face's edges, whatever the order
        cr = np.cross(X[:, f[2]] - X[:, f[1]], X[:, f[0]] - X[:, f[1]])
        np.testing.assert_allclose(S[:, j], 0.5 * np.linalg.norm(cr, axis=1), rtol=1e-12)
    # the volumes are |det| / 6 and the divergence theorem over the boundary gives their sum
    det = np.abs(np.linalg.det(X[:, 1:] - X[:, :1]))
    np.testing.assert_allclose(vol, det / 6, rtol=1e-10)
    h = np.arange(4 * n)
    inner = nb >= 0
    assert (nb[nb[inner]] == h[inner]).all()
    assert ((nb[inner] >> 2) != (h[inner] >> 2)).all()
    flat_n = nrm.reshape(-1, 3)
    assert np.abs(flat_n[inner] + flat_n[nb[inner]]).max() <= 1e-12
    np.testing.assert_allclose(ar[inner], ar[nb[inner]], rtol=1e-12)
    for hh in h[inner][:2000]:
        a = set(T[hh >> 2][list(G.VTK_TETRA_FACES[hh & 3])])
        k = nb[hh]
        assert a == set(T[k >> 2][list(G.VTK_TETRA_FACES[k & 3])])
    bnd = ~inner
    fcen = np.stack([X[:, list(f)].mean(1) for f in G.VTK_TETRA_FACES], 1).reshape(-1, 3)
    div = (fcen[bnd] * flat_n[bnd]).sum(1) @ ar[bnd] / 3.0
    np.testing.assert_allclose(div, vol.sum(), rtol=1e-10)
    return bnd, fcen


def test_oracle_pins_brick(brick):
    P, T, arrays = brick
    nb, nrm, ar, vol, cen = arrays
    bnd, fcen = _pins(P, T, arrays)
    np.testing.assert_allclose(vol.sum(), 1.0 * 0.05 * 0.05, rtol=1e-12)
    # boundary half-faces lie on the box surface; GEOMETRY_TYPE Cone codes by plane
    x = fcen[bnd]
    on = (np.isclose(x[:, 0], 0, atol=1e-15) | np.isclose(x[:, 0], 1, atol=1e-15)
          | np.isclose(x[:, 1], 0, atol=1e-15) | np.isclose(x[:, 1], 0.05, atol=1e-15)
          | np.isclose(x[:, 2], 0, atol=1e-15) | np.isclose(x[:, 2], 0.05, atol=1e-15))
    assert on.all()
    codes = nb[bnd]
    expect = np.where(x[:, 0] == 0.0, -3, np.where(x[:, 0] == 1.0, -5, -2))
    assert np.array_equal(codes, expect)
    # every half-face is inner or boundary, F = (4N + B) / 2
    n = T.shape[0]
    assert (4 * n + bnd.sum()) % 2 == 0
    assert (4 * n + bnd.sum()) // 2 == gold_brick()["n_faces"]


def test_oracle_pins_sphere(sphere):
    P, T, arrays = sphere
    bnd, fcen = _pins(P, T, arrays)
    assert (arrays[0][bnd] == -2).all()  # Cone rule: no face of a ball has normal +-x within 1e-7


def test_vtk_face_table_orientation():
    """vtkTetra's faces: each leaves out one point, and for a positively
    oriented tet the right-hand normal of each face points away from it."""
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1.0]])
    seen = set()
    for f in G.VTK_TETRA_FACES:
        opp = ({0, 1, 2, 3} - set(f)).pop()
        seen.add(opp)
        n = np.cross(X[f[1]] - X[f[0]], X[f[2]] - X[f[0]])
        assert n @ (X[opp] - X[f[0]]) < 0
    assert seen == {0, 1, 2, 3}


def test_orientation_polynomial_is_minus_triple_product():
    rng = np.random.default_rng(7)
    P = rng.normal(size=(4, 1000, 3))
    s = G.orientation_sum(*P)
    tp = (np.cross(P[1] - P[0], P[2] - P[0]) * (P[3] - P[0])).sum(1)
    np.testing.assert_allclose(s, -tp, rtol=1e-9, atol=1e-12)


def test_two_cells_by_hand():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1.0]])
    T = np.array([[0, 1, 2, 3], [1, 2, 3, 4]], np.int32)
    nb = G.neighbors(T)
    # cell 0 face 1 = points (1,2,3); in cell 1 those are local (0,1,2) = face 3 (0,2,1)
    expect = [-1, 7, -1, -1, -1, -1, -1, 1]
    assert nb.tolist() == expect
    g = M.tet_geometry(P, T, M.BOUND_FREE)
    assert g.neighbors.tolist() == expect
    m = M.bin_to_geo(g)
    assert m.n_faces == 7
    assert M.face_id_l(m.faces).tolist() == [0, 0, 0, 0, 1, 1, 1]
    assert M.face_id_r(m.faces).tolist() == [-1, 1, -1, -1, -1, -1, -1]
    assert m.cells["id_faces"].tolist() == [[0, 1, 2, 3], [4, 5, 6, 1]]
    assert m.cells["sign_n"].tolist() == [0b1111, 0b0111]
    assert M.cell_is_bound(m.cells).tolist() == [3, 3]  # 2-bit counter


def test_oracle_bin_to_geo_matches_reference(brick):
    _, _, (nb, nrm, ar, vol, cen) = brick
    faces, cells = G.convert_bin_to_geo(nb, nrm, ar, vol, cen)
    g = gold_brick()
    assert len(faces) == g["n_faces"]
    if not O.have_ref():
        pytest.skip("oracle/_ref not built (reference tree absent)")
    rc, rf = O.ref_convert_bin_to_geo(nb, nrm, ar, vol, cen)
    mc = np.frombuffer(rc.tobytes(), M.GEO_CELL_DTYPE)
    mf = np.frombuffer(rf.tobytes(), M.GEO_FACE_DTYPE)
    assert M.face_id_l(mf).tolist() == [f[0] for f in faces]
    assert M.face_id_r(mf).tolist() == [f[1] for f in faces]
    assert np.array_equal(mf["n"], np.array([f[2] for f in faces]))
    assert np.array_equal(mf["S"], np.array([f[3] for f in faces]))
    assert mc["id_faces"].tolist() == [c[0] for c in cells]
    assert mc["sign_n"].tolist() == [c[1] for c in cells]
    assert M.cell_is_bound(mc).tolist() == [c[2] for c in cells]


# ------------------------------------------------------- product vs oracle
@pytest.mark.parametrize("name", ["brick", "sphere"])
def test_reader_matches_oracle(vtk_files, name):
    P, T = G.read_legacy_vtk(vtk_files[name])
    p, t = M.read_vtk(vtk_files[name])
    assert np.array_equal(P.view(np.uint64), p.view(np.uint64))
    assert np.array_equal(T, t)


def test_geometry_bitwise_brick(vtk_files, brick):
    P, T, (nb, nrm, ar, vol, cen) = brick
    p, t = M.read_vtk(vtk_files["brick"])
    g = M.tet_geometry(p, t, M.BOUND_CONE)
    for a, b in ((nb, g.neighbors), (nrm, g.normals), (ar, g.areas), (vol, g.volumes), (cen, g.centers)):
        assert a.dtype == b.dtype and a.shape == b.shape and np.array_equal(a.view(np.uint8), b.view(np.uint8))
    gold = gold_brick()
    assert sha(g.neighbors) == gold["neighbors_sha256"] and sha(g.normals) == gold["normals_sha256"]
    assert sha(g.areas) == gold["areas_sha256"] and sha(g.volumes) == gold["volumes_sha256"]
    assert sha(g.centers) == gold["centers_sha256"]
    m = M.bin_to_geo(g)
    assert sha(m.cells) == gold["geo_cells_sha256"] and sha(m.faces) == gold["geo_faces_sha256"]
    if O.have_ref():
        rc, rf = O.ref_convert_bin_to_geo(g.neighbors, g.normals, g.areas, g.volumes, g.centers)
        assert np.array_equal(rc, m.cells.view(np.uint8).reshape(-1))
        assert np.array_equal(rf, m.faces.view(np.uint8).reshape(-1))


@pytest.mark.parametrize("rule", ["Cone", "Cube"])
def test_geometry_bitwise_sphere(vtk_files, sphere, rule):
    P, T, arrays = sphere
    if rule == "Cube":
        arrays = (G.set_type_of_bound(G.neighbors(T), arrays[1], "Cube"),) + tuple(arrays[1:])
    p, t = M.read_vtk(vtk_files["sphere"])
    g = M.tet_geometry(p, t, M.BOUND_CONE if rule == "Cone" else M.BOUND_FREE)
    for a, b in zip(arrays, (g.neighbors, g.normals, g.areas, g.volumes, g.centers)):
        assert np.array_equal(np.asarray(a).view(np.uint8), b.view(np.uint8))
    if O.have_ref():
        rc, rf = O.ref_convert_bin_to_geo(*arrays)
        m = M.bin_to_geo(g)
        assert np.array_equal(rc, m.cells.view(np.uint8).reshape(-1))
        assert np.array_equal(rf, m.faces.view(np.uint8).reshape(-1))


def test_build_geo_from_vtk_writes_reference_files(vtk_files, tmp_path):
    base = str(tmp_path) + os.sep
    g, m = M.build_geo_from_vtk(vtk_files["brick"], base)
    gold = gold_brick()
    n = gold["n_cells"]
    for name, count, rec in (("neighbors.bin", 4 * n, 4), ("normals.bin", n, 96), ("areas.bin", 4 * n, 8),
                             ("volume.bin", n, 8), ("centers.bin", n, 24), ("geo_cells.bin", n, 56),
                             ("geo_faces.bin", gold["n_faces"], 40)):
        raw = open(base + name, "rb").read()
        assert struct.unpack("<i", raw[:4])[0] == count and len(raw) == 4 + count * rec, name
    back = M.Mesh.read(base + "geo_faces.bin", base + "geo_cells.bin")
    assert sha(back.cells) == gold["geo_cells_sha256"] and sha(back.faces) == gold["geo_faces_sha256"]
    g2 = M.TetGeometry.read(base)
    assert sha(g2.normals) == gold["normals_sha256"] and sha(g2.centers) == gold["centers_sha256"]


# --------------------------------------------------------------- formats
def _write_vtk(path, P, T, version="3.0", binary=False, ptype="double", itype="vtktypeint64", extras=False):
    n, N = P.shape[0], T.shape[0]
    out = bytearray()

    def line(s):
        out.extend((s + "\n").encode())

    line(f"# vtk DataFile Version {version}")
    line("test mesh")
    line("BINARY" if binary else "ASCII")
    line("DATASET UNSTRUCTURED_GRID")
    if extras:
        line("FIELD FieldData 1")
        line("TIME 1 1 double")
        if binary:
            out.extend(struct.pack(">d", 0.25))
            line("")
        else:
            line("0.25")
    line(f"POINTS {n} {ptype}")
    if binary:
        out.extend(np.asarray(P, ">f8" if ptype == "double" else ">f4").tobytes())
        line("")
    else:
        vals = P if ptype == "double" else P.astype(np.float32)
        for row in vals:
            line(" ".join(repr(float(v)) if ptype == "double" else repr(np.float32(v).item()) for v in row))
    if extras and version.startswith("5"):
        line("METADATA")
        line("INFORMATION 0")
        line("")
    if version.startswith("5"):
        line(f"CELLS {N + 1} {4 * N}")
        off = np.arange(0, 4 * N + 1, 4)
        w = ">i8" if itype == "vtktypeint64" else ">i4"
        line(f"OFFSETS {itype}")
        if binary:
            out.extend(off.astype(w).tobytes())
            line("")
        else:
            line(" ".join(map(str, off)))
        line(f"CONNECTIVITY {itype}")
        if binary:
            out.extend(T.reshape(-1).astype(w).tobytes())
            line("")
        else:
            for row in T:
                line(" ".join(map(str, row)))
    else:
        line(f"CELLS {N} {5 * N}")
        rows = np.column_stack([np.full(N, 4), T]).astype(np.int64)
        if binary:
            out.extend(rows.reshape(-1).astype(">i4").tobytes())
            line("")
        else:
            for row in rows:
                line(" ".join(map(str, row)))
    line(f"CELL_TYPES {N}")
    if binary:
        out.extend(np.full(N, 10, ">i4").tobytes())
        line("")
    else:
        for _ in range(N):
            line("10")
    if extras:
        line(f"CELL_DATA {N}")
        line("SCALARS density double 1")
        line("LOOKUP_TABLE default")
        for _ in range(N):
            line("1.0")
    with open(path, "wb") as f:
        f.write(bytes(out))


@pytest.mark.parametrize("version", ["3.0", "4.2", "5.1"])
@pytest.mark.parametrize("binary", [False, True])
@pytest.mark.parametrize("ptype", ["double", "float"])
def test_reader_encodings(vtk_files, tmp_path, version, binary, ptype):
    P, T = M.read_vtk(vtk_files["sphere"])
    P, T = P[:400], None
    rng = np.random.default_rng(3)
    T = np.array([rng.choice(400, 4, replace=False) for _ in range(300)], np.int32)
    path = str(tmp_path / "m.vtk")
    _write_vtk(path, P, T, version, binary, ptype, itype="vtktypeint32" if binary else "vtktypeint64", extras=True)
    p, t = M.read_vtk(path)
    expect = P if ptype == "double" else P.astype(np.float32).astype(np.float64)
    assert np.array_equal(p, expect)
    assert np.array_equal(t, T)


def _expect_error(tmp_path, text, needle):
    path = tmp_path / "bad.vtk"
    path.write_text(text)
    with pytest.raises(ValueError, match=needle):
        M.read_vtk(str(path))


HEAD = "# vtk DataFile Version 3.0\nt\nASCII\nDATASET UNSTRUCTURED_GRID\n"
PTS = "POINTS 5 double\n0 0 0\n1 0 0\n0 1 0\n0 0 1\n1 1 1\n"


def test_reader_errors(tmp_path):
    _expect_error(tmp_path, "hello\n", "not a legacy VTK")
    _expect_error(tmp_path, "# vtk DataFile Version 3.0\nt\nASCII\nDATASET POLYDATA\n", "not UNSTRUCTURED_GRID")
    _expect_error(tmp_path, HEAD + PTS + "CELLS 1 5\n4 0 1 2 3\n", "no CELL_TYPES")
    _expect_error(tmp_path, HEAD + PTS + "CELLS 1 9\n8 0 1 2 3 4 0 1 2\nCELL_TYPES 1\n12\n", "VTK type 12")
    _expect_error(tmp_path, HEAD + PTS + "CELLS 1 5\n4 0 1 2 9\nCELL_TYPES 1\n10\n", "out of range")
    _expect_error(tmp_path, HEAD + "POINTS 5 double\n0 0 0\n1 0\n", "POINTS declares 15 values, more than")
    _expect_error(tmp_path, HEAD + "POINTS 2 double\n0 0 0\n1 0 x\n", "POINTS value")
    _expect_error(tmp_path, HEAD + PTS + "CELLS 2 10\n4 0 1 2 3\n", "CELLS")


def test_reader_rejects_counts_beyond_the_file(tmp_path):
    """Declared counts are checked against the bytes left before anything is sized from them:
    a corrupt count is an error message, not an allocation of its size."""
    big = 1 << 60
    _expect_error(tmp_path, HEAD + f"POINTS {big} double\n0 0 0\n", "more than the rest of the file")
    _expect_error(tmp_path, HEAD + f"POINTS {1 << 62} double\n0 0 0\n", "bad POINTS line")
    _expect_error(tmp_path, HEAD + PTS + f"CELLS 1 {big}\n4 0 1 2 3\n", "more than the rest of the file")
    _expect_error(tmp_path, HEAD + PTS + f"CELLS {big} 5\n4 0 1 2 3\n", "smaller than its cell count")
    _expect_error(tmp_path, HEAD + PTS + "CELLS 1 5\n4 0 1 2 3\n" + f"CELL_TYPES {big}\n10\n",
                  "more than the rest of the file")
    v5 = "# vtk DataFile Version 5.1\nt\nASCII\nDATASET UNSTRUCTURED_GRID\n"
    _expect_error(tmp_path, v5 + PTS + f"CELLS {big} 4\nOFFSETS vtktypeint64\n0 4\n", "OFFSETS declares")
    _expect_error(tmp_path, HEAD + PTS + f"FIELD f 1\na 3 {big} double\n1 2 3\n", "FIELD block is truncated")
    path = tmp_path / "bin.vtk"
    path.write_bytes(b"# vtk DataFile Version 3.0\nt\nBINARY\nDATASET UNSTRUCTURED_GRID\n"
                     + f"POINTS {1 << 40} double\n".encode() + bytes(48))
    with pytest.raises(ValueError, match="more than the rest of the file"):
        M.read_vtk(str(path))


def test_geometry_errors():
    P = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 1, 1.0], [-1, -1, -1.0]])
    with pytest.raises(ValueError, match="repeated"):
        M.tet_geometry(P, np.array([[0, 1, 1, 3]], np.int32))
    # three cells on the face (0, 1, 2)
    T = np.array([[0, 1, 2, 3], [0, 1, 2, 4], [0, 1, 2, 5]], np.int32)
    with pytest.raises(ValueError, match="more than one cell"):
        M.tet_geometry(P, T)
    with pytest.raises(ValueError, match="More than 1 neighbor"):
        G.neighbors(T)


# ------------------------------------------- the reference's test on brick.vtk
def test_brick_reference_test_oracle_run(vtk_files):
    """tests/src/off_rhllc_test.cpp:22-45 through the oracle step on the mesh
    library's geometry: bitwise the compiled reference's state at T = 0.5."""
    gold = gold_brick()
    p, t = M.read_vtk(vtk_files["brick"])
    m = M.bin_to_geo(M.tet_geometry(p, t, M.BOUND_CONE))
    W0 = np.tile(np.array(gold["W_init"]), (m.n_cells, 1))
    st = O.State(m, W0)
    st.ph.has_out_source = 1
    for k in range(5):
        st.ph.out_source_W[k] = gold["W_source"][k]
        st.ph.out_source_U[k] = gold["U_source"][k]
    t_end, taus, speeds = st.run(gold["cfl"], gold["h_min"], gold["steps"])
    assert sha(taus) == gold["taus_sha256"]
    assert t_end == gold["t"] and speeds[-1] == gold["last_speed"]
    assert sha(st.W) == gold["W_sha256"] and sha(st.U) == gold["U_sha256"]
    if O.have_ref():
        W_init, W_src, U_src = O.ref_cone_states()
        assert W_init.tolist() == gold["W_init"] and W_src.tolist() == gold["W_source"]
        assert U_src.tolist() == gold["U_source"]
