"""Host layout of the tiled step, checked against the mesh records (no GPU).

The tiled step (one workgroup per tile of 256 consecutive cells in the device
numbering) computes every face of a tile into LDS: the faces whose earlier
cell lies in the tile (a contiguous range, since faces are grouped by their
earlier cell) and the extra faces it shares with an earlier tile.  Each cell
then reads its four fluxes through 16-bit tile-local slots.  These tests
rebuild every cell's face list from the layout and compare it, slot by slot
and sign by sign, with the geo_cell_t records of the reference's format
(solvers_struct.h:120-154: id_faces, sign_n bit j = 1 when the cell is id_l).
"""
import numpy as np
import pytest

from radrhd_solver_amd import capi
from radrhd_solver_amd import mesh as M

T = 256


def _check(m, reorder):
    L = capi.tile_layout(m, reorder=reorder)
    N, F = m.n_cells, m.n_faces
    c_n2o, f_n2o = L["cell_new_to_old"], L["face_new_to_old"]
    assert np.array_equal(np.sort(c_n2o), np.arange(N))
    assert np.array_equal(np.sort(f_n2o), np.arange(F))
    meta, extra, slots = L["tile_meta"], L["tile_extra"], L["cell_slots"]
    nt = (N + T - 1) // T
    assert meta.shape == (nt, 4)
    # own ranges partition the faces in order; extras are listed tile by tile
    assert meta[0, 0] == 0 and (meta[1:, 0] == meta[:-1, 0] + meta[:-1, 1]).all()
    assert meta[-1, 0] + meta[-1, 1] == F
    assert meta[0, 2] == 0 and (meta[1:, 2] == meta[:-1, 2] + meta[:-1, 3]).all()
    assert meta[-1, 2] + meta[-1, 3] == extra.size
    assert L["max_slots"] == (meta[:, 1] + meta[:, 3]).max() <= 4 * T
    # the earlier / later cell of every device face, in the device numbering
    c_o2n = np.empty(N, np.int64)
    c_o2n[c_n2o] = np.arange(N)
    idl = M.face_id_l(m.faces)[f_n2o]
    idr = M.face_id_r(m.faces)[f_n2o]
    l = c_o2n[idl]
    r = np.where(idr >= 0, c_o2n[np.maximum(idr, 0)], -1)
    earlier = np.where(r >= 0, np.minimum(l, r), l)
    later = np.where(r >= 0, np.maximum(l, r), -1)
    tile_of_face = np.repeat(np.arange(nt), meta[:, 1])
    assert np.array_equal(earlier // T, tile_of_face)
    ex_tile = np.repeat(np.arange(nt), meta[:, 3])
    assert np.array_equal(later[extra] // T, ex_tile) and (earlier[extra] // T < ex_tile).all()
    # faces shared by two tiles are exactly the extras
    shared = (later >= 0) & (later // T != earlier // T)
    assert np.array_equal(np.sort(extra), np.nonzero(shared)[0])
    # every cell's slots name its own faces in slot order, with the sign bits
    ids = m.cells["id_faces"][c_n2o]
    sign = m.cells["sign_n"][c_n2o].astype(np.int64)
    tile = np.arange(N) // T
    ex = extra if extra.size else np.zeros(1, extra.dtype)
    for j in range(4):
        s = (slots[:, j // 2] >> (16 * (j % 2))) & 0xFFFF
        k = (s & 0x7FFF).astype(np.int64)
        own = meta[tile, 1]
        assert (k < own + meta[tile, 3]).all()
        dev_face = np.where(k < own, meta[tile, 0] + k, ex[np.clip(meta[tile, 2] + k - own, 0, ex.size - 1)])
        assert np.array_equal(f_n2o[dev_face], ids[:, j])
        assert np.array_equal((s & 0x8000) != 0, ((sign >> j) & 1) == 0)
    return L


@pytest.mark.parametrize("reorder", [True, False])
def test_tile_layout_matches_the_mesh_records(reorder):
    m = M.kuhn_box(12, 10, 9, seed=3)  # 6480 tets, 26 tiles (the last one partial)
    L = _check(m, reorder)
    assert L["tile_extra"].size > 0


def test_tile_layout_periodic_and_tiny_meshes():
    _check(M.kuhn_box(8, 8, 8, periodic=7, seed=5), True)  # faces across the box join far-apart tiles
    _check(M.kuhn_box(1, 1, 1, seed=5), True)  # 6 tets: one partial tile, no extras
    _check(M.kuhn_box(2, 1, 1, seed=5), False)


def test_tile_layout_on_the_reference_brick(tmp_path):
    """The reference's own test mesh (tests/data/brick.vtk, 118,783 tets, Cone boundaries)."""
    import lzma
    import os

    src = os.path.join(os.path.dirname(__file__), "golden", "brick.vtk.xz")
    path = tmp_path / "brick.vtk"
    with lzma.open(src, "rb") as f:
        path.write_bytes(f.read())
    P, Tt = M.read_vtk(str(path))
    m = M.bin_to_geo(M.tet_geometry(P, Tt, M.BOUND_CONE))
    _check(m, True)


def test_tile_layout_rejects_malformed_meshes():
    m = M.kuhn_box(2, 2, 2, seed=1)
    bad = M.Mesh(m.cells.copy(), m.faces.copy())
    bad.cells["sign_n"][0] ^= 1
    with pytest.raises(capi.RhllcError, match="disagrees"):
        capi.tile_layout(bad)
