"""Pins the Mie helper used by the physics sanity test: in the Rayleigh
limit (k a -> 0) the internal field of a sphere is uniform and equal to
3 / (eps + 2) times the incident field (Bohren & Huffman 5.12)."""
import numpy as np

import mie


def test_rayleigh_limit():
    pts = 1e-4 * np.array([[1.0, 2.0, 3.0], [-2.0, 1.0, 0.5], [0.0, -3.0, 2.0]])
    for eps in (4.0, 50.0 - 20.0j):
        m2 = eps
        E = mie.internal_field_bh(pts, 1.0, 0.001, np.sqrt(m2))
        np.testing.assert_allclose(E[:, 0], 3.0 / (m2 + 2.0), rtol=5e-3)
        assert np.abs(E[:, 1:]).max() < 5e-3


def test_vie_convention_is_conjugate():
    pts = np.array([[0.01, 0.02, -0.01]])
    e_bh = mie.internal_field_bh(pts, 6.2, 0.07, np.conj(np.sqrt(50 - 1j * 3.6)))
    e_v = mie.internal_field(pts, 6.2, 0.07, 50.0, 3.6, 1.0, 1.0)
    np.testing.assert_allclose(e_v, np.conj(e_bh), rtol=1e-12)
