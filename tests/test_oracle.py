"""Pins the CPU oracle (oracle/vie_oracle.py) against mathematics.

The reference repository ships no tests or golden vectors for this path, so
each oracle function is checked against an independent construction:
brute-force block-Toeplitz matrices (tests/bruteforce.py), closed forms,
algebraic identities, and numpy.linalg.solve for GMRES.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf

TINY = [(3, 4, 5), (2, 3, 2), (1, 5, 4), (2, 1, 1)]   # a single voxel along an axis: FFT length 2


def _vec(J):
    return J.reshape(-1, order="F")


def _field(v, dims, ql):
    return v.reshape(dims + (ql,), order="F")


def test_parity_tables_match_tensor_symmetry():
    gn, gk, bt = bf.parity_tables()
    np.testing.assert_array_equal(gn, vo.G_N_PAR)
    np.testing.assert_array_equal(gk, vo.G_K_PAR)
    np.testing.assert_array_equal(bt, vo.BT_PAR)


def test_nmp_matches_einsum():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((3, 4, 5)) + 1j * rng.standard_normal((3, 4, 5))
    B1 = rng.standard_normal((3, 7)) + 0j
    B2 = rng.standard_normal((4, 6)) + 0j
    B3 = rng.standard_normal((5, 2)) + 0j
    np.testing.assert_allclose(vo.nmp(A, B1, 1), np.einsum("ijk,im->mjk", A, B1), atol=1e-13)
    np.testing.assert_allclose(vo.nmp(A, B2, 2), np.einsum("ijk,jm->imk", A, B2), atol=1e-13)
    np.testing.assert_allclose(vo.nmp(A, B3, 3), np.einsum("ijk,km->ijm", A, B3), atol=1e-13)


def test_hosvd_full_rank_reconstructs():
    rng = np.random.default_rng(2)
    A = rng.standard_normal((4, 5, 6)) + 1j * rng.standard_normal((4, 5, 6))
    G, U1, U2, U3 = vo.hosvd(A)
    assert G.shape == (4, 5, 6)
    np.testing.assert_allclose(vo.hosvd_to_full(G, U1.T, U2.T, U3.T), A, atol=1e-12)


def test_hosvd_truncation_bound():
    # a smooth kernel is low rank; the truncation error obeys the HOSVD bound
    x = np.linspace(0.1, 1.0, 12)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    A = np.exp(-1j * 3.0 * np.sqrt(X ** 2 + Y ** 2 + Z ** 2)) / np.sqrt(X ** 2 + Y ** 2 + Z ** 2)
    tol = 1e-7
    G, U1, U2, U3 = vo.hosvd(A, tol)
    assert max(G.shape) < 12                      # actually truncated
    err = np.linalg.norm(vo.hosvd_to_full(G, U1.T, U2.T, U3.T) - A)
    # sum over modes of discarded singular values is bounded by 3 * n * (tol s1 / sqrt 3)
    s1 = np.linalg.svd(A.reshape(12, -1, order="F"), compute_uv=False)[0]
    assert err <= 3 * 12 * tol * s1 / np.sqrt(3)


@pytest.mark.parametrize("dims", TINY)
@pytest.mark.parametrize("op", ["N", "K"])
@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("form", ["dense", "tucker"])
def test_operator_matches_bruteforce_toeplitz(dims, op, pwl, form):
    rng = np.random.default_rng(10 + dims[0] + 3 * pwl + 7 * (op == "K"))
    A = bf.random_kernel(rng, dims, op, pwl)
    M = bf.dense_matrix(A, op, pwl)
    if form == "dense":
        T = vo.circulant_dense(A, pwl)
    else:
        T = (vo.fft_circ_tucker_pwl if pwl else vo.fft_circ_tucker_pwc)(A, None)
    ql = 12 if pwl else 3
    J = rng.standard_normal(dims + (ql,)) + 1j * rng.standard_normal(dims + (ql,))
    apply = vo.mvp_N if op == "N" else vo.mvp_K
    out = apply(J, T, pwl)
    ref = M @ _vec(J)
    np.testing.assert_allclose(_vec(out), ref, rtol=0, atol=1e-12 * np.abs(ref).max())
    # Hermitian adjoint = conjugate transpose of the same matrix
    outH = apply(J, T, pwl, adjoint=True)
    refH = M.conj().T @ _vec(J)
    np.testing.assert_allclose(_vec(outH), refH, rtol=0, atol=1e-12 * np.abs(refH).max())


def test_tucker_truncated_matches_dense_to_hosvd_level():
    rng = np.random.default_rng(5)
    dims = (6, 5, 7)
    # smooth (low-rank) kernel with the N parity structure
    A = np.zeros(dims + (6,), dtype=np.complex128)
    ii, jj, kk = np.meshgrid(*[np.arange(n) for n in dims], indexing="ij")
    R = np.sqrt(ii ** 2 + jj ** 2 + kk ** 2) + 1.0
    base = np.exp(-1j * 0.7 * R) / R
    odd = {0: (), 1: (0, 1), 2: (0, 2), 3: (), 4: (1, 2), 5: ()}
    for c in range(6):
        a = base.copy()
        for ax in odd[c]:
            a = a * [ii, jj, kk][ax]
        A[..., c] = a
    Td = vo.circulant_dense(A, False)
    Tt = vo.fft_circ_tucker_pwc(A, 1e-7)
    J = rng.standard_normal(dims + (3,)) + 1j * rng.standard_normal(dims + (3,))
    a = vo.mvp_N(J, Td, False)
    b = vo.mvp_N(J, Tt, False)
    assert np.linalg.norm(a - b) / np.linalg.norm(a) < 1e-6


@pytest.mark.parametrize("pwl", [False, True])
def test_gram_closed_form(pwl):
    res = 0.004
    g = vo.gram_diag(pwl, res)
    # moments of the unit cell: integral of 1 and of ((x - xc)/res)^2 over a cube of side res
    xs, ws = np.polynomial.legendre.leggauss(4)
    xs = xs * res / 2
    ws = ws * res / 2
    m0 = ws.sum() ** 3
    m2 = (ws * (xs / res) ** 2).sum() * ws.sum() ** 2
    want = np.tile([m0, m2, m2, m2], 3) if pwl else np.full(3, m0)
    np.testing.assert_allclose(g, want, rtol=1e-13)
    J = np.ones((2, 2, 2, len(g)), dtype=np.complex128)
    np.testing.assert_allclose(vo.mvp_invG(vo.mvp_G(J, res, pwl), res, pwl), J, rtol=1e-14)


@pytest.mark.parametrize("pwl", [False, True])
def test_mvp_vie_and_adjoint_against_dense(pwl):
    rng = np.random.default_rng(7 + pwl)
    dims = (3, 4, 3)
    A = bf.random_kernel(rng, dims, "N", pwl)
    M = bf.dense_matrix(A, "N", pwl)
    T = vo.circulant_dense(A, pwl)
    ql = 12 if pwl else 3
    nv = int(np.prod(dims))
    res = 0.005
    eps = 1 + 50 * rng.random(dims)
    sig = rng.random(dims)
    ce = vo.em_constants(298e6)["ce"]
    mat = vo.material(eps, sig, ce)
    idxS = np.sort(rng.choice(nv, size=nv // 2, replace=False))
    isol = vo.idx_solution(idxS, nv, ql)
    # dense VIE matrix: P^T (I - D invG M) P
    D = np.tile(_vec(mat["Mcr"]), ql)
    ig = np.repeat(1.0 / vo.gram_diag(pwl, res), nv)
    Afull = np.eye(ql * nv) - (D * ig)[:, None] * M
    Avie = Afull[np.ix_(isol, isol)]
    x = rng.standard_normal(len(isol)) + 1j * rng.standard_normal(len(isol))
    y = vo.mvp_vie(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    np.testing.assert_allclose(y, Avie @ x, atol=1e-12 * np.abs(y).max())
    yh = vo.mvp_vie_adjoint(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    np.testing.assert_allclose(yh, Avie.conj().T @ x, atol=1e-12 * np.abs(yh).max())


def test_gmres_solves_dense_system():
    rng = np.random.default_rng(3)
    n = 40
    A = np.eye(n) + 0.3 * (rng.standard_normal((n, n)) + 1j * rng.standard_normal((n, n))) / np.sqrt(n)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    x, it, rv = vo.gmres_vie(lambda v: A @ v, b, restart=10, tol=1e-12, maxit=50)
    np.testing.assert_allclose(x, np.linalg.solve(A, b), atol=1e-10)
    assert rv[-1] <= 1e-12
    # iteration bookkeeping: resvec has 1 + 10*(outit-1) + p entries
    p, outit = it
    assert len(rv) == 1 + 10 * (outit - 1) + p
    # recurrence residual equals the true residual up to rounding
    assert abs(np.linalg.norm(b - A @ x) / np.linalg.norm(b) - rv[-1]) < 1e-10


def test_gmres_happy_path_identity_is_one_step():
    b = np.arange(1, 6).astype(np.complex128)
    # A = 2 I: the Krylov space is invariant after one step; H(2,1) = 0 is never
    # divided by before the residual test, as the reference divides unguarded
    # (is_iter_gmres_vie.m:19-20) the case w = 0 would produce NaN; here w != 0
    x, it, rv = vo.gmres_vie(lambda v: 2.0 * v + 1e-3 * v[::-1], b, restart=5, tol=1e-10, maxit=3)
    np.testing.assert_allclose(2.0 * x + 1e-3 * x[::-1], b, atol=1e-9)


def test_gmres_zero_rhs_with_warm_start_is_noop():
    # when x0 already solves the system to tolerance the outer loop never runs
    A = np.diag(np.arange(1, 6).astype(np.complex128))
    b = np.ones(5, dtype=np.complex128)
    x0 = np.linalg.solve(A, b)
    x, it, rv = vo.gmres_vie(lambda v: A @ v, b, restart=5, tol=1e-6, maxit=3, x0=x0)
    assert it.tolist() == [1, 0] and len(rv) == 1
    np.testing.assert_array_equal(x, x0)


def test_rhs_vie_formula():
    ce = vo.em_constants(128e6)["ce"]
    eps = np.full((2, 2, 1), 40.0)
    sig = np.full((2, 2, 1), 0.5)
    mat = vo.material(eps, sig, ce)
    idxS = np.array([0, 3])
    Ue = np.ones(6, dtype=np.complex128)
    r = vo.rhs_vie(ce, mat["Mr"], mat["Mc"], idxS, 3, Ue)
    Mr = 40 + 0.5 / ce
    np.testing.assert_allclose(r, np.full(6, ce * (Mr - 1) / Mr), rtol=1e-15)
