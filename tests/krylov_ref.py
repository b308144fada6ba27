"""numpy restatement of the device COCR solver (vie_cocr) for the tests.

Not part of the oracle: the reference solves with GMRES only.  This is the
same recurrence as marie-tools_amd/csrc/vie_cocr.hip in FP64 (conjugate
orthogonal conjugate residual in the bilinear form [u, v] = sum u v w,
Chronopoulos-Gear arrangement with one reduction per iteration, the
residual replaced by b - A x at convergence), so that the FP64 device solver
can be compared iterate by iterate.  tests/test_cocr_ref.py pins it against
dense solves and checks the symmetry that makes the recurrence apply.
"""
from __future__ import annotations

import numpy as np

from oracle import vie_oracle as vo


def vie_weights(Mr, Mc, idxS, ql, res, pwl, adjoint=False):
    """Weights of the bilinear form in which mvp_vie (adjoint: its conjugate
    transpose) is symmetric: 1/d_i (adjoint: conj(d_i)) with
    d_i = (Mc/Mr)_s invG_c for packed unknown i = c nS + s; 0 where d_i = 0."""
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[idxS]
    invg = 1.0 / vo.gram_diag(pwl, res)
    d = np.concatenate([mcr * g for g in invg[:ql]])
    if adjoint:
        return np.conj(d)
    w = np.zeros_like(d)
    nz = d != 0
    w[nz] = 1.0 / d[nz]
    return w


def bil(u, v, w):
    return np.sum(u * v * w)


def cocr(afun, b, w, tol, maxit, x0=None):
    """Returns (x, resvec, nmatvec): resvec = ||r||/||b|| at the start, after
    every iteration, and after the final replacement r = b - A x."""
    bn = np.linalg.norm(b)
    nmv = 0
    if x0 is not None and np.count_nonzero(x0):
        x = x0.astype(np.complex128).copy()
        r = b - afun(x)
        nmv += 1
    else:
        x = np.zeros_like(b, dtype=np.complex128)
        r = b.astype(np.complex128).copy()
    res = [np.linalg.norm(r) / bn]
    xs = np.zeros_like(x)
    p = Ap = None
    rho = sig = None
    fresh = True
    its = 0
    while True:
        if np.linalg.norm(r) < tol * bn or its >= maxit:
            x = x + xs
            xs[:] = 0
            r = b - afun(x)
            nmv += 1
            res.append(np.linalg.norm(r) / bn)
            if np.linalg.norm(r) < tol * bn or its >= maxit:
                break
        Ar = afun(r)
        nmv += 1
        its += 1
        rr, aa = bil(r, Ar, w), bil(Ar, Ar, w)
        if fresh:
            beta, sig = 0.0, aa
            p, Ap = r.copy(), Ar.copy()
        else:
            beta = rr / rho
            sig = aa + 2 * beta * bil(Ar, Ap, w) + beta * beta * sig
            p = r + beta * p
            Ap = Ar + beta * Ap
        alpha = rr / sig
        rho = rr
        fresh = False
        xs = xs + alpha * p
        r = r - alpha * Ap
        res.append(np.linalg.norm(r) / bn)
    return x, np.array(res), nmv
