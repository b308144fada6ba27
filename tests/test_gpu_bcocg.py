"""GPU tests of the device block COCG solver (vie_bcocg, through the C ABI).

* FP64 block COCG equals the numpy recurrence of tests/krylov_ref.py
  (pinned against dense solves in test_bcocg_ref.py) iterate by iterate:
  resvec to 1e-6 relative, the same operator applications, X to 1e-9; the
  FP64 residual of every column (the oracle's operator) meets tol.
* The adjoint system with its own weights.
* Mixed form: every FP64 residual (oracle operator) below tol, X within
  10 tol of the FP64 solution.
* More rhs than one group (32 + 4): every column meets tol and equals its
  single COCR solve within 10 tol; dependent rhs (a breakdown of the block
  recurrence) finish through the COCR fallback.
* One rhs (a one-column block) equals the numpy recurrence; maxit ends with
  the FP64 residual as the last resvec entry; a warm start at the solution
  spends one FP64 application and returns it unchanged.
"""
import numpy as np
import pytest

import krylov_ref as kr
from test_gpu_cocr import _system, _op, _w, _rel

pytestmark = pytest.mark.gpu


def _relres(s, X, adjoint=False):
    f = s["adj"] if adjoint else s["fwd"]
    return max(np.linalg.norm(s["B"][:, j] - f(X[:, j])) / np.linalg.norm(s["B"][:, j])
               for j in range(X.shape[1]))


@pytest.mark.parametrize("pwl", [False, True])
def test_fp64_block_cocg_matches_numpy_recurrence(pwl):
    print(f"case pwl={pwl}", flush=True)
    s = _system(81 + pwl, (6, 5, 7), pwl, nrhs=4)
    op = _op(s)
    tol = 1e-10
    X, info = op.block_cocg(s["B"], tol=tol, maxit=400)
    Xr, rvr, nmvr, brk = kr.block_cocg(s["fwd"], s["B"], _w(s), tol, 400)
    assert not brk
    for j in range(4):
        assert info["nmatvec"][j] == nmvr[j]
        assert len(info["resvec"][j]) == len(rvr[j])
        np.testing.assert_allclose(info["resvec"][j], rvr[j], rtol=1e-6, atol=1e-14)
    assert _rel(X, Xr) < 1e-9
    assert _relres(s, X) < tol
    op.close()


def test_fp64_block_cocg_adjoint():
    s = _system(83, (6, 5, 7), False, nrhs=3)
    op = _op(s)
    tol = 1e-10
    X, info = op.block_cocg(s["B"], tol=tol, maxit=400, adjoint=True)
    Xr, rvr, nmvr, brk = kr.block_cocg(s["adj"], s["B"], _w(s, adjoint=True), tol, 400)
    assert not brk and info["nmatvec"][0] == nmvr[0]
    assert _rel(X, Xr) < 1e-9
    assert _relres(s, X, adjoint=True) < tol
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
def test_mixed_block_cocg_meets_fp64_residual(pwl):
    print(f"case pwl={pwl}", flush=True)
    s = _system(85 + pwl, (6, 5, 7), pwl, nrhs=4)
    op = _op(s)
    tol = 1e-8
    X, info = op.block_cocg(s["B"], tol=tol, maxit=600, precision="mixed")
    assert _relres(s, X) < tol                                          # the oracle's FP64 operator
    assert min(info["nreplace"]) >= 1 and all(r[-1] < tol for r in info["resvec"])
    X64, _ = op.block_cocg(s["B"], tol=1e-12, maxit=600)
    assert _rel(X, X64) < 10 * tol
    op.close()


def test_block_cocg_two_groups_and_dependent_rhs():
    s = _system(87, (6, 5, 7), True, nrhs=36)
    op = _op(s)
    tol = 1e-9
    X, info = op.block_cocg(s["B"], tol=tol, maxit=400)
    assert _relres(s, X) < tol
    for j in (0, 17, 31, 32, 35):                                       # both groups
        xj, _ = op.cocr(s["B"][:, j], tol=1e-12, maxit=400)
        assert _rel(X[:, j], xj) < 10 * tol
    # dependent rhs: the block recurrence breaks down, COCR finishes the group
    s3 = _system(89, (6, 5, 7), False, nrhs=3)
    s3["B"][:, 2] = 2.0 * s3["B"][:, 0]
    op3 = _op(s3)
    X3, info3 = op3.block_cocg(s3["B"], tol=tol, maxit=400)
    assert _relres(s3, X3) < tol
    assert np.abs(X3[:, 2] - 2.0 * X3[:, 0]).max() < 10 * tol * np.abs(X3[:, 2]).max()
    op.close()
    op3.close()


def test_block_cocg_one_rhs_maxit_and_warm_start():
    s = _system(91, (6, 5, 7), False, nrhs=2)
    op = _op(s)
    b = s["B"][:, :1]
    # one rhs (a 1-column block): the numpy recurrence, iterate by iterate
    x, info = op.block_cocg(np.asfortranarray(b), tol=1e-10, maxit=400)
    xr, rvr, nmvr, brk = kr.block_cocg(s["fwd"], b, _w(s), 1e-10, 400)
    assert not brk and info["nmatvec"][0] == nmvr[0]
    np.testing.assert_allclose(info["resvec"][0], rvr[0], rtol=1e-6, atol=1e-14)
    assert _rel(x, xr) < 1e-9
    # maxit: three iterations, then the FP64 residual as the last entry
    xm, infm = op.block_cocg(s["B"], tol=1e-14, maxit=3)
    for j in range(2):
        assert len(infm["resvec"][j]) == 5 and infm["nmatvec"][j] == 4
        true = np.linalg.norm(s["B"][:, j] - s["fwd"](xm[:, j])) / np.linalg.norm(s["B"][:, j])
        assert abs(infm["resvec"][j][-1] - true) < 1e-12
    # warm start at the solution: no iteration, one FP64 residual
    X64, _ = op.block_cocg(s["B"], tol=1e-12, maxit=400)
    Xw, infw = op.block_cocg(s["B"], x0=X64, tol=1e-10, maxit=400)
    assert infw["nmatvec"] == [1, 1] and _rel(Xw, X64) == 0.0
    op.close()
