"""CPU restatement of the reference's mesh preprocessing (TEST INFRASTRUCTURE ONLY).

Only tests/ may import this module, as the checker of the mesh library's
legacy-VTK reader, geometry arrays and geo records.  It follows, step for
step and in the same floating-point order (numpy elementwise IEEE double, no
fused operations):

* ``read_legacy_vtk``   -- the legacy ASCII format that VTK 9.3's
  vtkGenericDataObjectReader reads (lib/files_sys/include/reader_vtk.h:43-61);
  pure Python tokens, ``float()`` for coordinates.
* ``neighbors``         -- GetNeighborFace3D (lib/geometry/src/convert_vtk_geo.cpp:122-197):
  for every unvisited local face, the cells holding all three of its points
  (vtkUnstructuredGrid::GetCellNeighbors) and the matching local face of the
  one neighbour (GetNumberNeighborFace, any permutation of the triple).
* ``normals_areas``     -- NormalAndSquareFace3D (convert_vtk_geo.cpp:199-253),
  including its orientation polynomial ``sum`` verbatim in evaluation order.
* ``volumes``           -- GetVolume3D (convert_vtk_geo.cpp:270-307).
* ``centers``           -- CenterOfTetra3D (convert_vtk_geo.cpp:309-347).
* ``set_type_of_bound`` -- SetTypeOfBound for GEOMETRY_TYPE Cone and Cube
  (lib/geometry/src/make_internal_format.cpp:12-108; Cone is the build's
  type, include/solvers_config.h:29).
* ``convert_bin_to_geo``-- ConvertBinToGeo (lib/geometry/src/convert_bin_to_geo.cpp:3-48),
  a literal sequential loop with the -10 "processed" marker.

Third-party restatement: the local face order is VTK 9.3's vtkTetra face
table (VTK is not in the reference tree; its version is
cmake_modules.cmake:1): faces (0,1,3), (1,2,3), (2,0,3), (0,2,1).  The
reference's own files pin nothing about it beyond the orientation rule it
implies (outward normals for a positively oriented tet), which the tests check.

Eigen operation order used by the reference for 3-vectors: ``a.cross(b)`` by
components (y*z' - z*y', z*x' - x*z', x*y' - y*x'), ``dot`` and
``squaredNorm`` summed left to right, ``norm() = sqrt(squaredNorm())``,
``normalize()`` divides each component by that root when it is positive.
"""
from __future__ import annotations

import numpy as np

VTK_TETRA_FACES = ((0, 1, 3), (1, 2, 3), (2, 0, 3), (0, 2, 1))

E_NEIGH_UNDEF = -5  # lib/geometry/include/geo_types.h:143
E_NEIGH_OUT_BOUND = -1  # geo_types.h:145
E_BOUND_FREE, E_BOUND_LOCK, E_BOUND_OUT_SOURCE, E_BOUND_OUTER_SURFACE = -1, -2, -3, -5  # lib/global_types.h:25-31


def read_legacy_vtk(path):
    """(points[P,3], tets[N,4]) from a legacy ASCII VTK unstructured grid (version < 5)."""
    with open(path) as f:
        lines = f.read().split("\n")
    assert lines[0].startswith("# vtk DataFile Version")
    assert lines[2].strip() == "ASCII" and lines[3].split() == ["DATASET", "UNSTRUCTURED_GRID"]
    toks = " ".join(lines[4:]).split()
    i = 0
    pts = cells = types = None
    while i < len(toks):
        k = toks[i]
        if k == "POINTS":
            n, typ = int(toks[i + 1]), toks[i + 2]
            vals = [float(v) for v in toks[i + 3:i + 3 + 3 * n]]
            if typ == "float":
                vals = [float(np.float32(v)) for v in vals]
            pts = np.array(vals, dtype=np.float64).reshape(n, 3)
            i += 3 + 3 * n
        elif k == "CELLS":
            n, size = int(toks[i + 1]), int(toks[i + 2])
            flat = [int(v) for v in toks[i + 3:i + 3 + size]]
            cells, at = [], 0
            for _ in range(n):
                c = flat[at]
                cells.append(flat[at + 1:at + 1 + c])
                at += 1 + c
            i += 3 + size
        elif k == "CELL_TYPES":
            n = int(toks[i + 1])
            types = [int(v) for v in toks[i + 2:i + 2 + n]]
            i += 2 + n
        elif k in ("CELL_DATA", "POINT_DATA"):
            break
        else:
            raise ValueError(f"unexpected token {k}")
    assert all(t == 10 for t in types) and all(len(c) == 4 for c in cells)
    return pts, np.array(cells, dtype=np.int32)


def face_ids(tets, c, j):
    return [int(tets[c, q]) for q in VTK_TETRA_FACES[j]]


def neighbors(tets):
    """GetNeighborFace3D: int array [4N], neighbour cell*4 + face or -1."""
    n = tets.shape[0]
    point_cells = {}
    for c in range(n):
        for q in range(4):
            point_cells.setdefault(int(tets[c, q]), []).append(c)
    nb = [E_NEIGH_UNDEF] * (4 * n)
    for c in range(n):
        for j in range(4):
            if nb[4 * c + j] != E_NEIGH_UNDEF:
                continue
            a, b, d = face_ids(tets, c, j)
            # GetCellNeighbors: cells other than c that use all three points.
            cand = set(point_cells[a]) & set(point_cells[b]) & set(point_cells[d])
            cand.discard(c)
            if len(cand) == 1:
                k = cand.pop()
                match = [jj for jj in range(4) if sorted(face_ids(tets, k, jj)) == sorted((a, b, d))]
                if not match:
                    raise ValueError(f"neighbor {c} not found")
                nb[4 * c + j] = 4 * k + match[0]
                nb[4 * k + match[0]] = 4 * c + j
            elif len(cand) == 0:
                nb[4 * c + j] = E_NEIGH_OUT_BOUND
            else:
                raise ValueError("More than 1 neighbor")
    return np.array(nb, dtype=np.int32)


def _cross(a, b):
    return np.stack([a[..., 1] * b[..., 2] - a[..., 2] * b[..., 1],
                     a[..., 2] * b[..., 0] - a[..., 0] * b[..., 2],
                     a[..., 0] * b[..., 1] - a[..., 1] * b[..., 0]], axis=-1)


def _sqnorm(v):
    return (v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]


def normals_areas(points, tets):
    """NormalAndSquareFace3D for every (cell, face): normals[N,4,3], areas[4N]."""
    n = tets.shape[0]
    normals = np.zeros((n, 4, 3))
    areas = np.zeros((n, 4))
    for j, (i0, i1, i2) in enumerate(VTK_TETRA_FACES):
        P0, P1, P2 = points[tets[:, i0]], points[tets[:, i1]], points[tets[:, i2]]
        a, b = P1 - P0, P2 - P0
        nn = np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                       -a[:, 0] * b[:, 2] + a[:, 2] * b[:, 0],
                       a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=-1)
        z = _sqnorm(nn)
        areas[:, j] = 0.5 * np.sqrt(z)
        r = np.sqrt(z)
        unit = np.where((z > 0)[:, None], nn / np.where(r > 0, r, 1.0)[:, None], nn)
        # The cell point that is not on the face: the first in cell order.
        opp = np.array([[q for q in range(4) if q not in (i0, i1, i2)][0]] * n)
        P3 = points[tets[np.arange(n), opp]]
        s = orientation_sum(P0, P1, P2, P3)
        normals[:, j] = np.where((s < 0)[:, None], -unit, unit)
    return normals, areas.reshape(-1)


def orientation_sum(P0, P1, P2, P3):
    """The polynomial ``sum`` of convert_vtk_geo.cpp:241-245; the normal is
    turned when it is negative.  It equals -((P1-P0) x (P2-P0)) . (P3-P0)."""
    x0, y0, z0 = P0[..., 0], P0[..., 1], P0[..., 2]
    x1, y1, z1 = P1[..., 0], P1[..., 1], P1[..., 2]
    x2, y2, z2 = P2[..., 0], P2[..., 1], P2[..., 2]
    x3, y3, z3 = P3[..., 0], P3[..., 1], P3[..., 2]
    return (x1 * (y2 - y3) * z0 + x0 * (y3 - y2) * z1 + x0 * (y1 - y3) * z2 + z2 * (x1 * y3 - x1 * y0)
            + x3 * (z0 * (y1 - y2) + z1 * (y2 - y0) + z2 * (y0 - y1)) + z3 * (x1 * (y0 - y2) + x0 * (y2 - y1))
            + x2 * (z0 * (y3 - y1) + z1 * (y0 - y3) + z3 * (y1 - y0)))


def volumes(points, tets):
    P = points[tets]
    a, b, c = P[:, 1] - P[:, 0], P[:, 2] - P[:, 0], P[:, 3] - P[:, 0]
    bc = _cross(b, c)
    return -((a[:, 0] * bc[:, 0] + a[:, 1] * bc[:, 1]) + a[:, 2] * bc[:, 2]) / 6.0


def centers(points, tets):
    P = points[tets]

    def area(q0, q1, q2):
        return 0.5 * np.sqrt(_sqnorm(_cross(q1 - q0, q2 - q0)))

    S = [area(P[:, 1], P[:, 2], P[:, 3]), area(P[:, 0], P[:, 2], P[:, 3]),
         area(P[:, 0], P[:, 1], P[:, 3]), area(P[:, 0], P[:, 1], P[:, 2])]
    tot = ((S[0] + S[1]) + S[2]) + S[3]
    num = ((S[0][:, None] * P[:, 0] + S[1][:, None] * P[:, 1]) + S[2][:, None] * P[:, 2]) + S[3][:, None] * P[:, 3]
    return num / tot[:, None]


def set_type_of_bound(nb, normals, geometry="Cone"):
    nb = nb.copy()
    n = normals.reshape(-1, 3)
    for h in np.nonzero(nb < 0)[0]:
        if geometry == "Cube":
            nb[h] = E_BOUND_FREE
            continue
        d_minus = n[h] - np.array([-1.0, 0.0, 0.0])
        d_plus = n[h] - np.array([1.0, 0.0, 0.0])
        if np.sqrt(_sqnorm(d_minus)) < 1e-7:
            nb[h] = E_BOUND_OUT_SOURCE
        elif np.sqrt(_sqnorm(d_plus)) < 1e-7:
            nb[h] = E_BOUND_OUTER_SURFACE
        else:
            nb[h] = E_BOUND_LOCK
    return nb


def build_data_from_vtk(points, tets, geometry="Cone"):
    """The five arrays: neighbors[4N], normals[N,4,3], areas[4N], volumes[N], centers[N,3]."""
    nb = neighbors(tets)
    nrm, ar = normals_areas(points, tets)
    return set_type_of_bound(nb, nrm, geometry), nrm, ar, volumes(points, tets), centers(points, tets)


def convert_bin_to_geo(nb, normals, areas, vols, cents):
    """Faces as (id_l, id_r, n[3], S) tuples and cells as (id_faces[4], sign bits, is_bound, V, center)."""
    nb = [int(v) for v in nb]
    N = len(vols)
    cell_ids = [[0] * 4 for _ in range(N)]
    sign = [0] * N
    is_bound = [0] * N
    faces = []
    cc = 0
    for i in range(4 * N):
        idx = nb[i]
        if idx == -10:
            continue
        c, j = divmod(i, 4)
        sign[c] |= 1 << j
        cell_ids[c][j] = cc
        nb[i] = -10
        if idx >= 0:
            k, jj = divmod(idx, 4)
            sign[k] &= ~(1 << jj)
            cell_ids[k][jj] = cc
            nb[idx] = -10
            id_r = k
        else:
            id_r = idx
            is_bound[c] = (is_bound[c] + 1) & 3  # 2-bit field
        faces.append((c, id_r, normals[c, j].copy(), float(areas[i])))
        cc += 1
    cells = [(cell_ids[c], sign[c], is_bound[c], float(vols[c]), cents[c].copy()) for c in range(N)]
    return faces, cells
