"""Direct block-Toeplitz construction of the VIE N/K operators on tiny grids.

This is an independent restatement used only to pin the FFT oracle: it builds
the dense matrix of the discrete convolution straight from the real-space
kernel A and the per-axis parity rule A(-m) = s * A(m), with no FFT, no
circulant embedding and no spectrum.  The coupling tables are spelled out
here as explicit dictionaries from the reference's MATLAB matrices
(mvp_N_pwl_tucker.m:12-13, mvp_K_pwc_tucker.m:12, assembly_fft_circ_tucker_*.m)
so that a wrong entry in the oracle's arrays fails the comparison.
"""
from __future__ import annotations

import itertools

import numpy as np

# N-tensor components xx xy xz yy yz zz -> 1..6 (G_N.m:14-19)
_N_COMP = {("x", "x"): 1, ("x", "y"): 2, ("x", "z"): 3, ("y", "x"): 2, ("y", "y"): 4,
           ("y", "z"): 5, ("z", "x"): 3, ("z", "y"): 5, ("z", "z"): 6}
# K coupling: out_p = sum_q sign * K_comp * J_q  (curl structure; mvp_K_pwc_tucker.m:12)
_K_COUP = {("x", "y"): (-1, 3), ("x", "z"): (+1, 2), ("y", "x"): (+1, 3),
           ("y", "z"): (-1, 1), ("z", "x"): (-1, 2), ("z", "y"): (+1, 1)}
# (l, lp) monomial pairs (VV_Nop.m:27-28): 1..10 for l <= lp; (lp, l) gets the sign
# of the odd moment, i.e. negative when exactly one of l, lp is the constant
_PAIRS = {(1, 1): 1, (2, 2): 2, (3, 3): 3, (4, 4): 4, (1, 2): 5, (1, 3): 6, (1, 4): 7,
          (2, 3): 8, (2, 4): 9, (3, 4): 10}

_AX = "xyz"


def _pair(l, lp):
    if (l, lp) in _PAIRS:
        return +1, _PAIRS[(l, lp)]
    i = _PAIRS[(lp, l)]
    sgn = -1 if (l == 1 or lp == 1) else +1
    return sgn, i


# parity per axis of each N component, from the symmetry of (d_i d_j + k^2 delta) g:
# odd along axis d iff exactly one of i, j is d
def _par_N(comp):
    i, j = {1: "xx", 2: "xy", 3: "xz", 4: "yy", 5: "yz", 6: "zz"}[comp]
    return tuple(-1 if ((i == a) != (j == a)) else +1 for a in _AX)


# K components are d_x g, d_y g, d_z g: odd along their own axis
def _par_K(comp):
    return tuple(-1 if _AX[comp - 1] == a else +1 for a in _AX)


# linear moments (x - xc): the pair (l, lp) is odd along axis a iff exactly one
# of l, lp is the monomial along a
def _par_pair(i):
    l, lp = [k for k, v in _PAIRS.items() if v == i][0]
    return tuple(-1 if ((l == a + 2) != (lp == a + 2)) else +1 for a in range(3))


def _ext(A3, s, m):
    """Kernel value at signed offset m from the stored octant A3[|m|]."""
    sign = 1
    for d in range(3):
        if m[d] < 0:
            sign *= s[d]
    return sign * A3[abs(m[0]), abs(m[1]), abs(m[2])]


def dense_matrix(A: np.ndarray, op: str, pwl: bool) -> np.ndarray:
    """Dense operator matrix on the n1 x n2 x n3 grid (Fortran voxel order,
    component-major unknowns), built voxel pair by voxel pair."""
    n1, n2, n3 = A.shape[:3]
    nv = n1 * n2 * n3
    ls = [1, 2, 3, 4] if pwl else [1]
    unk = [(p, l) for p in _AX for l in ls]
    ql = len(unk)
    M = np.zeros((ql * nv, ql * nv), dtype=np.complex128)
    vox = [(i, j, k) for k in range(n3) for j in range(n2) for i in range(n1)]
    for u, (p, l) in enumerate(unk):
        for v, (q, lp) in enumerate(unk):
            if op == "N":
                comp = _N_COMP[(p, q)]
                csign = 1
                par = _par_N(comp)
            else:
                if p == q:
                    continue
                csign, comp = _K_COUP[(p, q)]
                par = _par_K(comp)
            if pwl:
                psign, i = _pair(l, lp)
                ppar = _par_pair(i)
                A3 = A[:, :, :, i - 1, comp - 1]
                s = tuple(par[d] * ppar[d] for d in range(3))
                sign = csign * psign
            else:
                A3 = A[:, :, :, comp - 1]
                s = par
                sign = csign
            for a, r in enumerate(vox):
                for b, rp in enumerate(vox):
                    m = (r[0] - rp[0], r[1] - rp[1], r[2] - rp[2])
                    M[u * nv + a, v * nv + b] = sign * _ext(A3, s, m)
    return M


def parity_tables():
    """Parity tables derived here from the tensor symmetries, for comparison
    with the hard-coded reference tables."""
    gn = np.array([[_par_N(c)[d] for c in range(1, 7)] for d in range(3)])
    gk = np.array([[_par_K(c)[d] for c in range(1, 4)] for d in range(3)])
    bt = np.array([[_par_pair(i)[d] for i in range(1, 11)] for d in range(3)])
    return gn, gk, bt


def random_kernel(rng, dims, op: str, pwl: bool) -> np.ndarray:
    """Random complex real-space kernel that respects the parity rule
    (odd components vanish on their zero-offset planes)."""
    n1, n2, n3 = dims
    ncomp = 6 if op == "N" else 3
    par = _par_N if op == "N" else _par_K
    if pwl:
        A = rng.standard_normal((n1, n2, n3, 10, ncomp)) + 1j * rng.standard_normal((n1, n2, n3, 10, ncomp))
        for i, c in itertools.product(range(10), range(ncomp)):
            s = tuple(par(c + 1)[d] * _par_pair(i + 1)[d] for d in range(3))
            _zero_odd(A[:, :, :, i, c], s)
    else:
        A = rng.standard_normal((n1, n2, n3, ncomp)) + 1j * rng.standard_normal((n1, n2, n3, ncomp))
        for c in range(ncomp):
            _zero_odd(A[:, :, :, c], par(c + 1))
    return A


def _zero_odd(A3, s):
    if s[0] < 0:
        A3[0, :, :] = 0
    if s[1] < 0:
        A3[:, 0, :] = 0
    if s[2] < 0:
        A3[:, :, 0] = 0
