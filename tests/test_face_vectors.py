"""The mixed path's 16-byte face vectors (rhllc_pack_face_vectors, the form
rhllc_create uploads): the FP32 vector plus three signed 7-bit multiples of
Q = 2^(E - 30) in the fourth word (bits 0-20; the tiled step keeps the face's
tile-local cell above them).  The decode below is the kernels'
unpack_sn written in numpy: Q from the FP32 components by bit operations,
k Q + fp32 in FP64.  Pinned by what the format promises: the rebuilt vector
is within 2^-30 of the largest component, the FP64 sum that rebuilds it is
exact, and the four faces of a tetrahedron, which close exactly in real
arithmetic, close to that bound instead of FP32's 2^-24."""
from fractions import Fraction

import numpy as np
import pytest

from radrhd_solver_amd import capi

SEED = 20261202


def _decode(p):
    """unpack_sn (rhllc_kernels.hip) in numpy."""
    h = p[:, :3].astype(np.float32)
    m = np.abs(h).max(axis=1)
    pw = (m.view(np.uint32) & np.uint32(0x7F800000)).view(np.float32).astype(np.float64)
    Q = pw * 2.0 ** -30
    w = p[:, 3].view(np.int32)
    k = [((w << s) >> 25).astype(np.float64) for s in (25, 18, 11)]
    return np.stack([k[i] * Q + h[:, i].astype(np.float64) for i in range(3)], axis=1), k, Q


def _vectors(rng, n):
    scale = 10.0 ** rng.uniform(-8, 8, n)
    v = rng.uniform(-1, 1, (n, 3)) * scale[:, None]
    v[: n // 8, 0] *= 1e-9  # near the y-z plane
    v[n // 8: n // 4, :2] = 0.0  # on an axis
    return v


def test_rebuilt_vectors_are_within_2_pow_33():
    rng = np.random.default_rng(SEED)
    e = _vectors(rng, 200_000)
    p = capi.pack_face_vectors(e)
    assert np.array_equal(p[:, :3], e.astype(np.float32))  # the first three words are the FP32 vector
    s, _, _ = _decode(p)
    m = np.abs(e).max(axis=1)
    err = np.abs(s - e).max(axis=1) / m
    fp32 = np.abs(p[:, :3].astype(np.float64) - e).max(axis=1) / m
    print("max error relative to the largest component", err.max(), "FP32 alone", fp32.max())
    assert err.max() <= 2.0 ** -30
    assert fp32.max() > 2.0 ** -26  # the residual bits do restore what FP32 loses
    assert (p[:, 3].view(np.uint32) >> 21 == 0).all()  # no tile-local cell outside a layout


def test_the_rebuilding_sum_is_exact():
    rng = np.random.default_rng(SEED + 1)
    e = _vectors(rng, 2000)
    p = capi.pack_face_vectors(e)
    s, k, Q = _decode(p)
    for i in range(0, 2000, 7):
        for c in range(3):
            exact = Fraction(float(p[i, c])) + Fraction(float(k[c][i])) * Fraction(float(Q[i]))
            assert Fraction(float(s[i, c])) == exact


def test_tetrahedron_faces_close_to_the_bound():
    rng = np.random.default_rng(SEED + 2)
    worst, worst32 = 0.0, 0.0
    for _ in range(500):
        x = rng.uniform(-1, 1, (4, 3)) * 10.0 ** rng.uniform(-4, 2)
        # outward S n of the four faces; they sum to zero in exact arithmetic
        faces = [(1, 2, 3, 0), (0, 3, 2, 1), (0, 1, 3, 2), (0, 2, 1, 3)]
        e = []
        for a, b, c, opp in faces:
            s = 0.5 * np.cross(x[b] - x[a], x[c] - x[a])
            if np.dot(s, x[a] - x[opp]) < 0:
                s = -s
            e.append(s)
        e = np.array(e)
        size = np.abs(e).max()
        p = capi.pack_face_vectors(e)
        s, _, _ = _decode(p)
        ref = e.sum(axis=0)  # the FP64 geometry's own closure (rounding of e)
        worst = max(worst, np.abs(s.sum(axis=0) - ref).max() / size)
        worst32 = max(worst32, np.abs(p[:, :3].astype(np.float64).sum(axis=0) - ref).max() / size)
    print("closure relative to the largest face component: packed", worst, "FP32", worst32)
    assert worst <= 4 * 2.0 ** -30
    assert worst32 > 10 * worst


def test_zero_vector_and_empty_input():
    p = capi.pack_face_vectors(np.zeros((1, 3)))
    assert np.array_equal(p, np.zeros((1, 4), np.float32))
    assert capi.pack_face_vectors(np.zeros((0, 3))).shape == (0, 4)
    s, _, _ = _decode(p)
    assert np.array_equal(s, np.zeros((1, 3)))
