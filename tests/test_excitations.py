"""Incident fields and their projection on the voxel basis (marie_vie.phantoms).

* Projection: closed forms (polynomial fields and the plane wave, exact
  cube integrals), and the reference's moment convention: the coupling
  kernels average over the cube with the Gauss weights over 8 and take the
  linear moments with 0.5 * node (Assemble_rwg_coupling_matrix_N_x.cpp:118-144),
  scaled by res^3 (sSVD_rwgBasis.m:55-79), i.e. the moment <T_l, E>; the VIE
  system works on coefficients, invG of the moments (mvp_invG_pwl.m:3-7).
* Hertzian dipole: the closed-form field equals (k^2 + grad div)(g p),
  g = exp(-j k R) / (4 pi R) (the e^{+j omega t} convention of G_N.m:4),
  evaluated by finite differences of g p, and it is divergence free and
  solves the vector Helmholtz equation away from the source.
* Surface-current patterns: sums of such dipoles on the enclosing cylinder;
  the same Maxwell checks, and linearity in the element phases.
"""
import math

import numpy as np
import pytest

from marie_vie import phantoms


def _tiny(dims=(3, 2, 2), res=0.01, origin=(0.013, -0.02, 0.031)):
    eps = np.full(dims, 2.0)
    sig = np.zeros(dims)
    idxS = np.arange(int(np.prod(dims)))
    return phantoms.Phantom(dims, res, eps, sig, idxS, origin=origin)


def _centres(ph):
    cx, cy, cz = ph.centres()
    n1, n2 = ph.dims[0], ph.dims[1]
    s = ph.idxS
    return cx[s % n1], cy[(s // n1) % n2], cz[s // (n1 * n2)]


def test_projection_of_polynomials_is_exact():
    ph = _tiny()
    h = ph.res
    X, Y, Z = _centres(ph)
    nS = len(ph.idxS)
    # E = (x^2, y z, 3): PWC coefficients = cube averages
    f = lambda x, y, z: (x ** 2, y * z, np.full(np.shape(x), 3.0 + 0j))
    c = phantoms.project(ph, f, False).reshape(3, nS)
    np.testing.assert_allclose(c[0], X ** 2 + h ** 2 / 12, rtol=1e-13)
    np.testing.assert_allclose(c[1], Y * Z, rtol=1e-13)
    np.testing.assert_allclose(c[2], 3.0, rtol=1e-14)
    # PWL: x^2 = (xc + h xi)^2 = xc^2 + 2 xc h xi + h^2 xi^2 -> const xc^2 + h^2/12,
    # x-linear 2 xc h (coefficient of xi, = invG of the moment)
    c = phantoms.project(ph, f, True).reshape(3, 4, nS)
    np.testing.assert_allclose(c[0, 0], X ** 2 + h ** 2 / 12, rtol=1e-13)
    np.testing.assert_allclose(c[0, 1], 2 * X * h, rtol=1e-12)
    np.testing.assert_allclose(c[0, 2:], 0, atol=1e-18)
    # y z = (yc + h eta)(zc + h zeta): y-linear zc h, z-linear yc h
    np.testing.assert_allclose(c[1, 0], Y * Z, rtol=1e-13)
    np.testing.assert_allclose(c[1, 2], Z * h, rtol=1e-12)
    np.testing.assert_allclose(c[1, 3], Y * h, rtol=1e-12)
    np.testing.assert_allclose(c[1, 1], 0, atol=1e-18)


def test_projection_of_the_plane_wave_matches_exact_integrals():
    ph = _tiny(res=0.02)
    k0 = phantoms.constants(298e6)["k0"] * 7.0          # a short wavelength makes the moments large
    h = ph.res
    X, Y, Z = _centres(ph)
    nS = len(ph.idxS)
    c = phantoms.project(ph, phantoms.plane_wave(k0), True, nq=8).reshape(3, 4, nS)
    a = k0 * h / 2
    avg = np.sin(a) / a                                  # average of exp(-j k h zeta) over zeta in [-1/2, 1/2]
    # int_{-1/2}^{1/2} zeta e^{-j b zeta} d zeta = j d/db (sin a / a) = j (a cos a - sin a) / (2 a^2),
    # b = k h, a = b / 2
    mom = 1j * (a * math.cos(a) - math.sin(a)) / (2 * a * a)
    e = np.exp(-1j * k0 * Z)
    np.testing.assert_allclose(c[0, 0], e * avg, rtol=1e-12)
    np.testing.assert_allclose(c[0, 3], 12 * e * mom, rtol=1e-10)
    np.testing.assert_allclose(c[0, 1:3], 0, atol=1e-15)
    np.testing.assert_allclose(c[1:], 0, atol=0)


def test_projection_is_invG_of_the_reference_moments():
    """Moments as the reference's coupling kernels form them: weight
    w_i w_j w_k / 8, linear moment 0.5 * node, times res^3; the coefficients
    the VIE system uses are invG of those moments (res^3 and res^3 / 12)."""
    ph = _tiny(res=0.004)
    nq = 3
    k0 = phantoms.constants(128e6)["k0"]
    field = phantoms.dipoles(k0, 1, 0.3, 0.1, seed=5)[0]
    node, wq = np.polynomial.legendre.leggauss(nq)
    X, Y, Z = _centres(ph)
    h = ph.res
    nS = len(ph.idxS)
    mom = np.zeros((3, 4, nS), complex)
    for i in range(nq):
        for j in range(nq):
            for k in range(nq):
                weight = wq[i] * wq[j] * wq[k] / 8.0
                E = field(X + node[i] * h / 2, Y + node[j] * h / 2, Z + node[k] * h / 2)
                lin = [1.0, 0.5 * node[i], 0.5 * node[j], 0.5 * node[k]]
                for p in range(3):
                    for l in range(4):
                        mom[p, l] += weight * lin[l] * E[p]
    mom *= h ** 3
    gram = np.array([h ** 3, h ** 3 / 12, h ** 3 / 12, h ** 3 / 12])
    want = (mom / gram[None, :, None]).reshape(-1)
    got = phantoms.project(ph, field, True, nq=nq)
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def _fd_grad_div(F, x, d):
    """(grad div F)(x) by central differences of step d (F: R^3 -> C^3)."""
    e = np.eye(3) * d
    out = np.zeros(3, complex)
    for i in range(3):
        for j in range(3):
            # d_i d_j F_j
            if i == j:
                out[i] += (F(x + e[i])[j] - 2 * F(x)[j] + F(x - e[i])[j]) / d ** 2
            else:
                out[i] += (F(x + e[i] + e[j])[j] - F(x + e[i] - e[j])[j] - F(x - e[i] + e[j])[j]
                           + F(x - e[i] - e[j])[j]) / (4 * d ** 2)
    return out


def _fd_div(F, x, d):
    e = np.eye(3) * d
    return sum((F(x + e[i])[i] - F(x - e[i])[i]) / (2 * d) for i in range(3))


def _fd_laplacian(F, x, d):
    e = np.eye(3) * d
    return sum((F(x + e[i]) - 2 * F(x) + F(x - e[i])) / d ** 2 for i in range(3))


def _vec(field):
    return lambda x: np.array(field(np.array(x[0]), np.array(x[1]), np.array(x[2])), dtype=complex)


@pytest.mark.parametrize("seed", [2004, 2005])
def test_dipole_field_is_the_dyadic_green_function(seed):
    k0 = phantoms.constants(298e6)["k0"]
    f = phantoms.dipoles(k0, 3, 0.30, 0.15, seed=seed)
    rng = np.random.default_rng(seed)
    for field in f:
        E = _vec(field)
        # recover the dipole position and moment direction from the closure
        pos, pdir = field.__defaults__
        g = lambda x: np.exp(-1j * k0 * np.linalg.norm(x - pos)) / (4 * math.pi * np.linalg.norm(x - pos))
        A = lambda x: g(x) * pdir                     # g p
        for _ in range(3):
            x = rng.uniform(-0.1, 0.1, 3)
            want = k0 ** 2 * A(x) + _fd_grad_div(A, x, 2e-4)
            got = E(x)
            assert np.abs(got - want).max() <= 1e-5 * np.abs(got).max()
            # div E = 0 and (laplacian + k^2) E = 0 away from the source
            assert abs(_fd_div(E, x, 1e-4)) <= 1e-6 * np.abs(got).max() / 0.05
            lap = _fd_laplacian(E, x, 4e-4) + k0 ** 2 * got
            assert np.abs(lap).max() <= 1e-3 * k0 ** 2 * np.abs(got).max()


def test_surface_patterns_are_source_free_and_linear_in_the_phases():
    k0 = phantoms.constants(123e6)["k0"]
    pats = phantoms.surface_patterns(k0, 2, 0.30, 0.30, 6, seed=2005)
    rng = np.random.default_rng(9)
    for field in pats:
        E = _vec(field)
        elems, ph = field.__defaults__
        assert len(elems) == 6 and np.allclose(np.abs(ph), 1.0)
        for e in elems:
            pos, pdir = e.__defaults__
            assert abs(math.hypot(pos[0], pos[1]) - 0.30) < 1e-12 and abs(pos[2]) <= 0.30
            assert abs(np.dot(pdir, [pos[0], pos[1], 0.0])) < 1e-12     # tangential to the cylinder
        for _ in range(2):
            x = rng.uniform(-0.1, 0.1, 3)
            got = E(x)
            want = sum(c * np.array(e(*x), dtype=complex) for e, c in zip(elems, ph))
            np.testing.assert_allclose(got, want, rtol=1e-13)
            assert abs(_fd_div(E, x, 1e-4)) <= 1e-6 * np.abs(got).max() / 0.05
            lap = _fd_laplacian(E, x, 4e-4) + k0 ** 2 * got
            assert np.abs(lap).max() <= 1e-3 * k0 ** 2 * np.abs(got).max()
