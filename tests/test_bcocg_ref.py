"""CPU tests of tests/krylov_ref.block_cocg, the numpy restatement of the
device block COCG solver (vie_bcocg), on dense random-kernel VIE systems of
the oracle (no counterpart in the reference, whose solver is GMRES):

* it solves a block of rhs to 1e-12 and matches numpy's dense solves, for
  the forward and the adjoint operator, PWC and PWL;
* its block Krylov space is larger than a single rhs's, so it needs fewer
  iterations than COCR needs for the slowest rhs alone;
* the residual norms it reports from C are those of B - A X (Q orthonormal);
* linearly dependent rhs are a breakdown, reported (the device solver then
  finishes with COCR); a warm start and maxit behave as in COCR.
"""
import numpy as np
import pytest

import krylov_ref as kr
from test_cocr_ref import _system, _dense


def _block(s, nrhs, seed):
    rng = np.random.default_rng(seed)
    n = len(s["b"])
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    B[:, 0] = s["b"]
    return B


def _w(s, adjoint=False):
    return kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], s["pwl"], adjoint)


@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("adjoint", [False, True])
def test_block_cocg_solves_dense_systems(pwl, adjoint):
    s = _system(41 + pwl, (4, 4, 4), pwl, frac=0.8)
    f = s["adj"] if adjoint else s["fwd"]
    n = len(s["b"])
    M = _dense(f, n)
    B = _block(s, 4, 7)
    X, rv, nmv, brk = kr.block_cocg(f, B, _w(s, adjoint), 1e-12, 400)
    assert not brk
    Xd = np.linalg.solve(M, B)
    for j in range(4):
        assert rv[j][-1] < 1e-12
        assert np.abs(X[:, j] - Xd[:, j]).max() / np.abs(Xd[:, j]).max() < 1e-10
        assert nmv[j] == len(rv[j]) - 1               # one per iteration, one final replacement


def test_block_cocg_fewer_iterations_than_cocr():
    s = _system(43, (4, 4, 4), False, frac=0.8)
    w = _w(s)
    B = _block(s, 6, 9)
    X, rv, nmv, brk = kr.block_cocg(s["fwd"], B, w, 1e-10, 600)
    assert not brk
    single = [kr.cocr(s["fwd"], B[:, j], w, 1e-10, 600)[2] for j in range(6)]
    assert nmv[0] < 0.8 * max(single)


def test_block_cocg_recurrence_norms_are_true_norms():
    """||C e_j|| (the reported recurrence residual) tracks ||b_j - A x_j||
    while the solve is far from rounding level: stopped after k iterations,
    the last recurrence entry and the replacement entry agree."""
    s = _system(45, (3, 4, 3), True)
    B = _block(s, 3, 11)
    for k in (3, 7):
        X, rv, _, _ = kr.block_cocg(s["fwd"], B, _w(s), 1e-14, k)
        for j in range(3):
            true = np.linalg.norm(B[:, j] - s["fwd"](X[:, j])) / np.linalg.norm(B[:, j])
            assert abs(rv[j][-1] - true) < 1e-12
            assert abs(rv[j][-2] - true) < 1e-8 * max(true, 1e-3)
            assert len(rv[j]) == k + 2


def test_block_cocg_dependent_rhs_break_down_and_warm_start():
    s = _system(47, (3, 4, 3), False)
    w = _w(s)
    B = _block(s, 3, 13)
    B[:, 2] = 2.0 * B[:, 0]
    _, _, _, brk = kr.block_cocg(s["fwd"], B, w, 1e-10, 100)
    assert brk
    # a block Krylov space that exhausts the 63 unknowns (4 rhs, 16
    # iterations) leaves dependent columns in Q': the recurrence restarts
    # from B - A X and still converges
    M = _dense(s["fwd"], len(s["b"]))
    B4 = _block(s, 4, 7)
    X4, rv, _, brk = kr.block_cocg(s["fwd"], B4, w, 1e-13, 400)
    assert not brk and len(rv[0]) - 2 > 16 and rv[0][-1] < 1e-13
    X4d = np.linalg.solve(M, B4)
    assert np.abs(X4 - X4d).max() / np.abs(X4d).max() < 1e-11
    B = _block(s, 3, 13)
    M = _dense(s["fwd"], len(s["b"]))
    Xd = np.linalg.solve(M, B)
    X0 = Xd + 1e-4 * np.random.default_rng(3).standard_normal(Xd.shape)
    X, rv, nmv, brk = kr.block_cocg(s["fwd"], B, w, 1e-12, 400, X0=X0)
    assert not brk and rv[0][0] < 1e-2
    assert np.abs(X - Xd).max() / np.abs(Xd).max() < 1e-10
    assert nmv[0] == len(rv[0])                       # + the initial residual of X0
