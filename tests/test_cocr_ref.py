"""CPU checks of the COCR solver's premise and of its numpy restatement.

* The packed VIE operator A = I - D N (mvp_vie.m) is symmetric in the
  bilinear form [u, v] = sum u_i v_i / d_i, and its adjoint in
  sum u_i v_i conj(d_i): checked on the oracle's dense matrices, whose
  N is pinned by brute-force Toeplitz products (test_oracle.py).  The
  unweighted form is not symmetric, so a solver with the wrong weights
  fails here.
* tests/krylov_ref.cocr (the recurrence of vie_cocr.hip) solves those
  systems to the dense solution.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf
import krylov_ref as kr


def _system(seed, dims, pwl, frac=0.6, res=0.005, scale=0.6, zero_contrast=False):
    rng = np.random.default_rng(seed)
    A = bf.random_kernel(rng, dims, "N", pwl)
    nv = int(np.prod(dims))
    eps = 1 + 60 * rng.random(dims)
    sig = rng.random(dims)
    idxS = np.sort(rng.choice(nv, size=max(1, int(frac * nv)), replace=False))
    if zero_contrast:          # one scatterer voxel of vacuum: Mc = 0 there
        v = np.unravel_index(idxS[0], dims, order="F")
        eps[v], sig[v] = 1.0, 0.0
    mat = vo.material(eps, sig, vo.em_constants(298e6)["ce"])
    ql = 12 if pwl else 3
    isol = vo.idx_solution(idxS, nv, ql)
    n = len(isol)
    b = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    if zero_contrast:          # b = ce (Mc/Mr) E_inc (ie_solver_vie.m:18) vanishes there
        b[np.arange(ql) * len(idxS)] = 0
    T0 = vo.circulant_dense(A, pwl)
    d = b - vo.mvp_vie(b, mat["Mr"], mat["Mc"], isol, dims, res, T0, pwl)
    A = A * (scale * np.linalg.norm(b) / np.linalg.norm(d))
    T = vo.circulant_dense(A, pwl)
    fwd = lambda v: vo.mvp_vie(v, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    adj = lambda v: vo.mvp_vie_adjoint(v, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    return dict(A=A, mat=mat, idxS=idxS, isol=isol, b=b, fwd=fwd, adj=adj, ql=ql, res=res, pwl=pwl, dims=dims)


def _dense(f, n):
    E = np.eye(n, dtype=np.complex128)
    return np.stack([f(E[:, k]) for k in range(n)], axis=1)


@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("adjoint", [False, True])
def test_vie_operator_symmetric_in_weighted_form(pwl, adjoint):
    s = _system(11 + pwl, (3, 4, 3), pwl)
    n = len(s["b"])
    M = _dense(s["adj"] if adjoint else s["fwd"], n)
    w = kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], pwl, adjoint)
    WM = w[:, None] * M
    asym = np.abs(WM - WM.T).max() / np.abs(WM).max()
    assert asym < 1e-12
    # the plain bilinear form (no weights) is not: the weights carry the contrast and invG
    assert np.abs(M - M.T).max() / np.abs(M).max() > 1e-3
    # nor the forward weights for the adjoint (or vice versa)
    w2 = kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], pwl, not adjoint)
    WM2 = w2[:, None] * M
    assert np.abs(WM2 - WM2.T).max() / np.abs(WM2).max() > 1e-3


@pytest.mark.parametrize("pwl", [False, True])
def test_cocr_ref_solves_dense_system(pwl):
    s = _system(21 + pwl, (3, 4, 3), pwl)
    n = len(s["b"])
    M = _dense(s["fwd"], n)
    xd = np.linalg.solve(M, s["b"])
    w = kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], pwl)
    x, rv, nmv = kr.cocr(s["fwd"], s["b"], w, 1e-12, 500)
    assert rv[-1] < 1e-12
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-10
    assert nmv == len(rv) - 1                    # one per iteration, one final replacement
    # adjoint system with its own weights
    Mh = _dense(s["adj"], n)
    assert np.abs(Mh - M.conj().T).max() < 1e-12 * np.abs(M).max()
    wh = kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], pwl, adjoint=True)
    xh, rvh, _ = kr.cocr(s["adj"], s["b"], wh, 1e-12, 500)
    xhd = np.linalg.solve(Mh, s["b"])
    assert np.abs(xh - xhd).max() / np.abs(xhd).max() < 1e-10


def test_cocr_ref_warm_start_zero_contrast_and_maxit():
    s = _system(31, (3, 4, 3), False, zero_contrast=True)
    n = len(s["b"])
    M = _dense(s["fwd"], n)
    w = kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], False)
    assert np.count_nonzero(w == 0) == 3           # the vacuum voxel's three unknowns
    xd = np.linalg.solve(M, s["b"])
    assert np.abs(xd[np.arange(3) * len(s["idxS"])]).max() == 0   # they stay zero
    x, rv, _ = kr.cocr(s["fwd"], s["b"], w, 1e-12, 500)
    assert np.abs(x - xd).max() / np.abs(xd).max() < 1e-10
    x0 = xd + 1e-3 * (np.random.default_rng(5).standard_normal(n) + 0j)
    xw, rvw, nmvw = kr.cocr(s["fwd"], s["b"], w, 1e-12, 500, x0=x0)
    assert rvw[0] < 1e-2 and np.abs(xw - xd).max() / np.abs(xd).max() < 1e-10
    assert nmvw == len(rvw)                         # + the initial residual of x0
    # maxit stops with the replacement residual as the last entry
    xm, rvm, nmvm = kr.cocr(s["fwd"], s["b"], w, 1e-14, 3)
    assert len(rvm) == 5 and nmvm == 4
    assert abs(rvm[-1] - np.linalg.norm(s["b"] - M @ xm) / np.linalg.norm(s["b"])) < 1e-12
