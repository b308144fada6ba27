"""The near-field Green tensor against the reference's own formulation.

tests/golden/near_field_{N,K}.npz hold the N- and K-operator entries at the
offsets {0,1}^3 as the reference computes them: the Gauss-theorem reduction
to face pairs (surface_surface_kernels_Nop.m, surface_surface_kernels_Kop.m)
with the reference's DIRECTFN C++ for the touching pairs, restated in
oracle/near_field.py and written by tests/golden/make_near_field.py (5 mm /
298 MHz PWC and PWL, 4 mm / 128 MHz PWL).  The library's near field
(marie_vie.green: integration by parts onto the faces, Duffy boxes at the
singular corner; its own scheme) must agree with it to quadrature precision.

The oracle's face-pair classification is pinned by the adjacency tables of
points_mapping.m (held below as data); where the compiled DIRECTFN exists
(only beside the reference checkout) the oracle is also checked against a
closed form and against the golden values themselves.
"""
import itertools
import math
import os

import numpy as np
import pytest

from oracle import near_field as nf

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = np.load(os.path.join(HERE, "golden", "near_field_N.npz"))
GOLD_K = np.load(os.path.join(HERE, "golden", "near_field_K.npz"))
OFFSETS = list(itertools.product((0, 1), repeat=3))

# adjacency of observation face (row) and source face (column), faces
# [-x +x -y +y -z +z], 0 none, 1 ST, 2 EA, 3 VA: points_mapping.m:62-326
_, S, E, V = 0, 1, 2, 3
ADJ = {
    (1, 1, 1): [[S, 0, E, E, E, E], [0, S, E, E, E, E], [E, E, S, 0, E, E], [E, E, 0, S, E, E],
                [E, E, E, E, S, 0], [E, E, E, E, 0, S]],
    (2, 1, 1): [[0, S, E, E, E, E], [0] * 6, [0, E, E, 0, V, V], [0, E, 0, E, V, V],
                [0, E, V, V, E, 0], [0, E, V, V, 0, E]],
    (2, 1, 2): [[0, E, V, V, 0, E], [0] * 6, [0, V, V, 0, 0, V], [0, V, 0, V, 0, V],
                [0, E, V, V, 0, E], [0] * 6],
    (1, 1, 2): [[E, 0, V, V, 0, E], [0, E, V, V, 0, E], [V, V, E, 0, 0, E], [V, V, 0, E, 0, E],
                [E, E, E, E, 0, S], [0] * 6],
    (1, 2, 1): [[E, 0, 0, E, V, V], [0, E, 0, E, V, V], [E, E, 0, S, E, E], [0] * 6,
                [V, V, 0, E, E, 0], [V, V, 0, E, 0, E]],
    (2, 2, 1): [[0, E, 0, E, V, V], [0] * 6, [0, E, 0, E, V, V], [0] * 6,
                [0, V, 0, V, V, 0], [0, V, 0, V, 0, V]],
    (1, 2, 2): [[V, 0, 0, V, 0, V], [0, V, 0, V, 0, V], [V, V, 0, E, 0, E], [0] * 6,
                [V, V, 0, E, 0, E], [0] * 6],
    (2, 2, 2): [[0, V, 0, V, 0, V], [0] * 6, [0, V, 0, V, 0, V], [0] * 6,
                [0, V, 0, V, 0, V], [0] * 6],
}


@pytest.mark.parametrize("m", sorted(ADJ))
def test_face_pair_classification_matches_points_mapping(m):
    got = [[nf.face_pair(m, 0.005, f, fp)[0] for fp in range(6)] for f in range(6)]
    assert got == ADJ[m]


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("case", ["c1_pwc", "c1_pwl", "c2_pwl"])
def test_library_near_field_matches_reference_formulation(case):
    from marie_vie import green
    res, k0, _, pwl = GOLD[case + "_params"]
    ref = GOLD[case + "_A"]
    got = green.near_field(float(res), float(k0), "N", bool(pwl), OFFSETS)
    for m in OFFSETS:
        e = _rel(got[m], ref[m])
        assert e < 1e-10, (case, m, e)
    # the golden entries are not trivial: self term, face, edge and corner neighbours differ
    mags = [np.abs(ref[m]).max() for m in OFFSETS]
    assert len({round(v, 16) for v in mags}) >= 4


# K: the reference's reduced kernels 2..4 carry (G - G0) / (j k0)^2, whose
# Gauss quadrature at Np = 15 (load_inputs.m:80) cancels digits when k0 dx is
# small: at 4 mm / 128 MHz (k0 dx = 0.011) the reference formulation's own
# values move by 1.9e-9 from 15 to 21 points, while the library's move by
# 8e-13 under a finer rule, so that case is held to 3e-9
K_TOL = {"c1_pwc": 1e-11, "c1_pwl": 1e-10, "c2_pwl": 3e-9}


@pytest.mark.parametrize("case", ["c1_pwc", "c1_pwl", "c2_pwl"])
def test_library_k_near_field_matches_reference_formulation(case):
    from marie_vie import green
    res, k0, _, pwl = GOLD_K[case + "_params"]
    ref = GOLD_K[case + "_A"]
    got = green.near_field(float(res), float(k0), "K", bool(pwl), OFFSETS)
    scale = np.abs(ref).max()       # the PWC self entry of K is exactly zero
    for m in OFFSETS:
        e = np.abs(got[m] - ref[m]).max() / scale
        assert e < K_TOL[case], (case, m, e)
    if not pwl:
        assert np.abs(ref[0, 0, 0]).max() == 0.0 and np.abs(got[(0, 0, 0)]).max() < 1e-15 * scale
    # antisymmetric tensor: a face neighbour along axis d couples only through the
    # component whose curl direction is d (grad g along the offset)
    for d, m in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
        off = [c for c in range(3) if c != d]
        assert np.abs(ref[m][0, off]).max() < 1e-12 * scale and np.abs(ref[m][0, d]) > 0.5 * scale


needs_dfn = pytest.mark.skipif(not os.path.exists(nf.LIB_PATH),
                               reason="oracle/_ref/libdirectfn.so is built only beside the reference checkout")


@needs_dfn
def test_directfn_static_self_term_closed_form():
    # int_S int_S 1 / (4 pi R) over a unit square = (4/3)(1 - sqrt 2 + 3 ln(1 + sqrt 2)) / (4 pi)
    typ, pts = nf.face_pair((1, 1, 1), 1.0, 4, 4)
    assert typ == nf.ST
    v = nf._singular(nf._lib(), typ, pts, 15, 1e-9, 1.0, nf.FACE_N[4], nf.FACE_N[4], 0, 0, 0)
    want = 4 / 3 * (1 - math.sqrt(2) + 3 * math.log(1 + math.sqrt(2))) / (4 * math.pi)
    assert abs(v.real - want) < 1e-9 * want


@needs_dfn
def test_oracle_reproduces_golden_entries():
    res, k0, Np, _ = GOLD["c1_pwc_params"]
    for m in ((0, 0, 0), (1, 1, 1)):
        got = nf.near_N(tuple(np.array(m) + 1), float(res), float(k0), 1, int(Np))
        np.testing.assert_allclose(got, GOLD["c1_pwc_A"][m], rtol=0, atol=1e-14 * np.abs(got).max())


@needs_dfn
def test_oracle_reproduces_golden_k_entries():
    res, k0, Np, _ = GOLD_K["c1_pwc_params"]
    for m in ((0, 0, 1), (1, 1, 1)):
        got = nf.near_K(tuple(np.array(m) + 1), float(res), float(k0), 1, int(Np))
        np.testing.assert_allclose(got, GOLD_K["c1_pwc_A"][m], rtol=0, atol=1e-14 * np.abs(got).max())
