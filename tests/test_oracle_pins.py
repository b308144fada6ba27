"""Known-answer tests for the oracle functions that no brute-force product pins.

The operator checks in test_oracle.py pass the outputs of ``em_constants``,
``material``, ``idx_solution`` and the HOSVD rank rule to both sides, so a
mistake in them would cancel there.  Each test below compares one of these
functions with a value fixed outside the oracle: published physical
constants, a hand-evaluated case of the MATLAB statement it restates, or a
tensor built with prescribed singular values.  Each is chosen so that a
plausible slip (a dropped 2*pi, the wrong time convention, voxel-major
packing, an exclusive rank, a missing sqrt(3) or a missing ``~= 0``) fails.
"""
import math

import numpy as np
import pytest

from oracle import vie_oracle as vo


# --------------------------------------------------------------------------
# em_constants (em_constants.m:119-139)
# --------------------------------------------------------------------------

def test_vacuum_permittivity_is_the_pre_2019_exact_value():
    # mu0 = 4 pi 1e-7 H/m and c0 = 299792458 m/s make eps0 = 1/(mu0 c0^2)
    # = 8.854187817620...e-12 F/m (CODATA 2014)
    assert vo.em_constants(298e6)["e0"] == pytest.approx(8.854187817620e-12, rel=1e-12)


@pytest.mark.parametrize("freq, k0", [(298e6, 6.2456), (128e6, 2.6827), (123e6, 2.5779)])
def test_wavenumber_at_the_workload_frequencies(freq, k0):
    # 7 T, 3 T and 2.89 T proton frequencies; k0 = 2 pi f / c0 rounded to 4 decimals
    emc = vo.em_constants(freq)
    assert emc["k0"] == pytest.approx(k0, abs=5e-5)
    # dispersion relation of free space: k0^2 = omega^2 mu0 eps0
    assert emc["k0"] ** 2 == pytest.approx(emc["omega"] ** 2 * vo.MU0 * emc["e0"], rel=1e-14)
    assert emc["omega"] == pytest.approx(2 * math.pi * freq, rel=1e-15)


def test_contrast_follows_the_exp_plus_j_omega_t_convention():
    # a tissue with eps_r 50 and sigma 0.6 S/m at 298 MHz has the complex
    # relative permittivity eps_r - j sigma / (omega eps0) = 50 - 36.19 j in
    # MARIE's e^{+j omega t} convention (G_N.m:4 uses e^{-j k0 R}); ce = j omega eps0
    emc = vo.em_constants(298e6)
    assert emc["ce"].real == 0.0
    assert emc["ce"].imag == pytest.approx(0.0165785, rel=1e-5)
    m = vo.material(np.array([50.0]), np.array([0.6]), emc["ce"])
    assert m["Mr"][0] == pytest.approx(50.0 - 36.1915j, abs=1e-4)
    assert m["Mc"][0] == pytest.approx(49.0 - 36.1915j, abs=1e-4)
    assert m["Mcr"][0] == pytest.approx((49.0 - 36.1915j) / (50.0 - 36.1915j), abs=1e-5)
    # air: Mr = 1, Mc = 0
    a = vo.material(np.array([1.0]), np.array([0.0]), emc["ce"])
    assert a["Mr"][0] == 1.0 and a["Mc"][0] == 0.0


# --------------------------------------------------------------------------
# idx_solution (geo_body_domain.m:24-35)
# --------------------------------------------------------------------------

def test_idx_solution_hand_case_pwc():
    # MATLAB: idxS = [2; 5] (1-based), nD = 6, ql = 3:
    # idx_solution((i-1)*nS+1 : i*nS) = (i-1)*nD + idxS  ->  [2 5 8 11 14 17]
    got = vo.idx_solution(np.array([1, 4]), 6, 3)
    np.testing.assert_array_equal(got, np.array([2, 5, 8, 11, 14, 17]) - 1)


def test_idx_solution_hand_case_pwl():
    # idxS = [1; 3] (1-based), nD = 3, ql = 12: block i holds (i-1)*3 + [1 3]
    got = vo.idx_solution(np.array([0, 2]), 3, 12)
    want = np.array([[3 * i + 1, 3 * i + 3] for i in range(12)]).reshape(-1) - 1
    np.testing.assert_array_equal(got, want)
    # component-major, not voxel-major: the first nS entries are component 1
    assert got[:2].tolist() == [0, 2] and got[2:4].tolist() == [3, 5]


def test_idx_solution_scatters_components_into_4d_slices():
    # Jp(idx_solution) = Jb (mvp_vie.m:12-13) puts unknown c of scatterer s at
    # Jp(idxS(s) + c nD): in the [n1,n2,n3,ql] field that is voxel idxS(s) of slice c
    dims, ql = (3, 2, 2), 4
    nD = int(np.prod(dims))
    idxS = np.array([1, 5, 10])
    Jb = np.arange(1, ql * len(idxS) + 1).astype(float)
    Jp = np.zeros(nD * ql)
    Jp[vo.idx_solution(idxS, nD, ql)] = Jb
    J4 = Jp.reshape(dims + (ql,), order="F")
    for c in range(ql):
        sl = J4[:, :, :, c].reshape(-1, order="F")
        np.testing.assert_array_equal(sl[idxS], Jb[c * 3:(c + 1) * 3])
        assert np.count_nonzero(sl) == 3


# --------------------------------------------------------------------------
# HOSVD rank rule (hosvd.m:13-25)
# --------------------------------------------------------------------------

def _tensor_with_mode_spectrum(sig, dims, seed):
    """sum_i sig_i u_i (x) v_i (x) w_i with orthonormal u, v, w: every mode
    unfolding then has exactly the singular values sig (and zeros)."""
    rng = np.random.default_rng(seed)
    r = len(sig)
    fac = []
    for n in dims:
        Q, _ = np.linalg.qr(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
        fac.append(Q)
    return np.einsum("i,ai,bi,ci->abc", np.asarray(sig, float), *fac)


def test_hosvd_rank_is_the_first_small_singular_value_inclusive():
    tol = 1e-7
    t = tol / math.sqrt(3.0)
    s1 = 7.0
    # sigma_3 lies between tol*s1/sqrt(3) and tol*s1, sigma_4 below the threshold
    sig = [s1, 2.1, s1 * 0.8 * tol, s1 * 0.5 * t, s1 * 0.1 * t]
    A = _tensor_with_mode_spectrum(sig, (6, 7, 8), 11)
    for d, n in enumerate(A.shape):
        Ad = np.moveaxis(A, d, 0).reshape(n, -1)
        np.testing.assert_allclose(np.linalg.svd(Ad, compute_uv=False)[:5], sig, rtol=1e-6, atol=1e-14)
    G, U1, U2, U3 = vo.hosvd(A, tol)
    # find(S <= tol*S(1)/sqrt(3) & S ~= 0, 1) = 4: the index of the first small value
    # is itself kept.  An exclusive rule gives 3; a threshold without sqrt(3) gives 3.
    assert G.shape == (4, 4, 4)
    assert U1.shape == (6, 4) and U2.shape == (7, 4) and U3.shape == (8, 4)


def test_hosvd_rank_ignores_exact_zeros():
    # a single nonzero entry: every unfolding has singular values [1, 0, 0, ...]
    # exactly; the "& S ~= 0" clause finds nothing, so r_d = n_d (hosvd.m:15-17)
    A = np.zeros((4, 5, 6), dtype=np.complex128)
    A[0, 0, 0] = 2.0
    G, U1, U2, U3 = vo.hosvd(A, 1e-7)
    assert G.shape == (4, 5, 6)
    np.testing.assert_allclose(vo.hosvd_to_full(G, U1.T, U2.T, U3.T), A, atol=1e-15)


def test_hosvd_without_tol_uses_the_numerical_rank():
    sig = [3.0, 1.0, 0.25]
    A = _tensor_with_mode_spectrum(sig, (5, 6, 7), 12)
    G, U1, U2, U3 = vo.hosvd(A)
    assert G.shape == (3, 3, 3)
    np.testing.assert_allclose(vo.hosvd_to_full(G, U1.T, U2.T, U3.T), A, atol=1e-13)


def test_hosvd_core_is_the_conjugate_projection():
    # G = A x1 U1^H x2 U2^H x3 U3^H (hosvd.m:35: nmp with conj(U)), so that
    # G x1 U1 x2 U2 x3 U3 reconstructs A; with U instead of conj(U) the
    # reconstruction of a complex tensor fails
    rng = np.random.default_rng(13)
    A = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    G, U1, U2, U3 = vo.hosvd(A, 1e-12)
    ref = np.einsum("abc,ai,bj,ck->ijk", A, U1.conj(), U2.conj(), U3.conj())
    np.testing.assert_allclose(G, ref, atol=1e-13)
