"""Real-space Green tensor of the VIE N and K operators (operator set-up).

assemble(...) returns A[n1, n2, n3, pwx, ncomp], the Galerkin integrals of
the N kernel of G_N.m (d_i d_j g off the diagonal, -d_j^2 g - d_k^2 g on it,
i.e. (d_i d_j + k0^2 delta_ij) g + delta_ij delta) or the K kernel d_j g,
g = exp(-i k0 R) / (4 pi R), between the voxel at the origin (basis) and the
voxel at offset m (test), with monomial weights T_l (test) and B_lp (basis).
It follows the three distance regimes of assembly_N.m:16-81 / assembly_K.m:17-80:
  far    : nq = 4 tensor Gauss rule over all offsets (HIP kernel, vie_green_quadrature)
  medium : nq = 8 rule over the first 8^3 offsets (the same kernel; order and
           extent per medium_rule)
  near   : offsets in {0,1}^3, where the kernel is singular.  The reference
           integrates these with surface-surface formulas and the DIRECTFN
           singular integrals; this module uses its own scheme instead:
           integrating by parts moves both derivatives onto the voxel
           boundaries and the monomials, leaving only the weakly singular g;
           every term then factorises per axis into a 1D weight of the
           relative coordinate u = r - r' (a polynomial on each interval
           [k h, (k+1) h]) times g(|u|), integrated by Gauss-Legendre on the
           elementary boxes with a Duffy transform on the boxes whose corner
           is the singular point u = 0.
The near-field values agree with the reference's surface-surface values
(DIRECTFN) at every near offset (tests/test_near_field.py), with the
far-field rule at separated offsets, and reproduce the static self term
+2 h^3/3 of the reference kernel, whose diagonal -d_j^2 g - d_k^2 g carries
the delta of -laplacian(g) (G_N.m:15-20; tests/test_green.py).  The set-up
runs once per operator, in FP64.
"""
from __future__ import annotations

import ctypes
import itertools
import math

import numpy as np

from . import _lib as L

# (l, lp) monomial pairs, 0-based (VV_Nop.m:27-28)
PAIRS = [(0, 0), (1, 1), (2, 2), (3, 3), (0, 1), (0, 2), (0, 3), (1, 2), (1, 3), (2, 3)]
# N tensor components xx xy xz yy yz zz (G_N.m:14-19)
N_COMPS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]

_G3 = np.polynomial.legendre.leggauss(3)


# ---------------------------------------------------------------------------
# per-axis weights of the relative coordinate u = x - x'
#   test cell  [m h - h/2, m h + h/2], test monomial  t(x)  = ((x - m h)/h)^td
#   basis cell [-h/2, h/2],            basis monomial b(x') = (x'/h)^bd
# kinds: ("V",)        both volumes: w(u) = int t(x) b(x - u) dx over the overlap
#        ("S", s)      test face x = m h + s h/2 (outward normal s), basis volume
#        ("P", s)      test volume, basis face x' = s h/2 (outward normal s)
#        ("SS", s, sp) both faces: point mass at u0 = m h + s h/2 - sp h/2
# ---------------------------------------------------------------------------
class Axis:
    def __init__(self, kind, m, td, bd, h):
        self.kind, self.m, self.td, self.bd, self.h = kind, m, td, bd, h

    def t(self, x):
        return ((x - self.m * self.h) / self.h) ** self.td

    def b(self, xp):
        return (xp / self.h) ** self.bd

    def support(self):
        """Elementary intervals k (u in [k h, (k+1) h]) or the point u0."""
        m, h = self.m, self.h
        k = self.kind
        if k[0] == "V":
            return [m - 1, m], None
        if k[0] == "S":
            return ([m] if k[1] > 0 else [m - 1]), None
        if k[0] == "P":
            return ([m - 1] if k[1] > 0 else [m]), None
        return None, m * h + k[1] * h / 2 - k[2] * h / 2

    def weight(self, u):
        m, h = self.m, self.h
        k = self.kind
        if k[0] == "V":
            lo = np.maximum(m * h - h / 2, u - h / 2)
            hi = np.minimum(m * h + h / 2, u + h / 2)
            ln = np.maximum(hi - lo, 0.0)
            x, w = _G3
            acc = 0.0
            for xi, wi in zip(x, w):
                xx = lo + (xi + 1) * ln / 2
                acc = acc + wi * self.t(xx) * self.b(xx - u)
            return acc * ln / 2
        if k[0] == "S":
            xf = m * h + k[1] * h / 2
            return k[1] * self.t(xf) * self.b(xf - u)
        if k[0] == "P":
            xpf = k[1] * h / 2
            return k[1] * self.t(xpf + u) * self.b(xpf)
        xf = m * h + k[1] * h / 2
        xpf = k[2] * h / 2
        return k[1] * k[2] * self.t(xf) * self.b(xpf)


class _Boxes:
    """Quadrature points of the elementary boxes with g(|u|) cached."""

    def __init__(self, h, k0, q_reg=8, q_sing=12):
        self.h, self.k0 = h, k0
        self.cache = {}
        self.xr, self.wr = np.polynomial.legendre.leggauss(q_reg)
        xs, ws = np.polynomial.legendre.leggauss(q_sing)
        self.xs, self.ws = (xs + 1) / 2, ws / 2          # on [0, 1]

    def g(self, R):
        return np.exp(-1j * self.k0 * R) / (4 * math.pi * R)

    def points(self, ks, fixed):
        """ks: per-axis elementary index or None; fixed: per-axis point value."""
        key = (tuple(ks), tuple(fixed))
        if key in self.cache:
            return self.cache[key]
        h = self.h
        cont = [d for d in range(3) if ks[d] is not None]
        D = len(cont)
        singular = all(ks[d] in (-1, 0) for d in cont) and all(
            fixed[d] == 0.0 for d in range(3) if ks[d] is None)
        if not singular:
            grids = [ks[d] * h + (self.xr + 1) * h / 2 for d in cont]
            wts = [self.wr * h / 2 for _ in cont]
            U = np.meshgrid(*grids, indexing="ij")
            Wm = np.ones_like(U[0])
            for a, w in enumerate(wts):
                sh = [1] * D
                sh[a] = -1
                Wm = Wm * w.reshape(sh)
            u = [None] * 3
            for a, d in enumerate(cont):
                u[d] = U[a].ravel()
            W = Wm.ravel()
        else:
            # Duffy: the box [0, h]^D (reflected per axis) split into D pyramids
            # with apex at the singular corner; v_p = h t, v_q = h t s_q
            sig = [(-1.0 if ks[d] == -1 else 1.0) for d in cont]
            us, Ws = [[] for _ in range(D)], []
            for p in range(D):
                others = [a for a in range(D) if a != p]
                grids = [self.xs] * D
                G = np.meshgrid(*grids, indexing="ij")
                wgt = np.ones_like(G[0])
                for a in range(D):
                    sh = [1] * D
                    sh[a] = -1
                    wgt = wgt * self.ws.reshape(sh)
                t = G[0]
                v = [None] * D
                v[p] = h * t
                for j, a in enumerate(others):
                    v[a] = h * t * G[j + 1]
                jac = h ** D * t ** (D - 1)
                for a in range(D):
                    us[a].append((sig[a] * v[a]).ravel())
                Ws.append((wgt * jac).ravel())
            u = [None] * 3
            for a, d in enumerate(cont):
                u[d] = np.concatenate(us[a])
            W = np.concatenate(Ws)
        npt = W.size
        for d in range(3):
            if u[d] is None:
                u[d] = np.full(npt, fixed[d])
        R = np.sqrt(u[0] ** 2 + u[1] ** 2 + u[2] ** 2)
        val = (u, W * self.g(R))
        self.cache[key] = val
        return val


def _integral(boxes: _Boxes, axes) -> complex:
    """int g(|u|) prod_d w_d(u_d) du over the product of the axis supports."""
    sup = [a.support() for a in axes]
    lists = [s[0] if s[0] is not None else [None] for s in sup]
    fixed = [s[1] if s[0] is None else 0.0 for s in sup]
    total = 0.0 + 0.0j
    for ks in itertools.product(*lists):
        u, Wg = boxes.points(list(ks), fixed)
        w = Wg
        for d in range(3):
            wd = axes[d].weight(u[d])
            w = w * wd
        total += w.sum()
    return total


def _mono(l):
    """Axis (0,1,2) of monomial l (1: x, 2: y, 3: z) or None for the constant."""
    return None if l == 0 else l - 1


def near_entry(boxes, m, l, lp, op, comp):
    """One near-field entry A(m, (l,lp), comp) by integration by parts."""
    h = boxes.h
    k0 = boxes.k0
    ta, ba = _mono(l), _mono(lp)
    td = [1 if ta == d else 0 for d in range(3)]
    bd = [1 if ba == d else 0 for d in range(3)]

    def ax(kinds, tdd, bdd):
        return [Axis(kinds[d], m[d], tdd[d], bdd[d], h) for d in range(3)]

    zero = [0, 0, 0]
    if op == "N":
        i, j = N_COMPS[comp]
        val = 0.0j
        # (1) - sum over test faces normal to i and basis faces normal to j
        for s in (1, -1):
            for sp in (1, -1):
                kinds = [("V",)] * 3
                if i == j:
                    kinds = [("SS", s, sp) if d == i else ("V",) for d in range(3)]
                else:
                    kinds = [("S", s) if d == i else (("P", sp) if d == j else ("V",)) for d in range(3)]
                val -= _integral(boxes, ax(kinds, td, bd))
        # (2) + test faces normal to i, d'_j B
        if ba == j:
            for s in (1, -1):
                kinds = [("S", s) if d == i else ("V",) for d in range(3)]
                val += _integral(boxes, ax(kinds, td, zero)) / h
        # (3) + d_i T, basis faces normal to j
        if ta == i:
            for sp in (1, -1):
                kinds = [("P", sp) if d == j else ("V",) for d in range(3)]
                val += _integral(boxes, ax(kinds, zero, bd)) / h
        # (4) - d_i T d'_j B
        if ta == i and ba == j:
            val -= _integral(boxes, ax([("V",)] * 3, zero, zero)) / h ** 2
        # (5) + k0^2 delta_ij T B
        if i == j:
            val += k0 ** 2 * _integral(boxes, ax([("V",)] * 3, td, bd))
        # (6) the reference kernel writes the diagonal as -d_j^2 g - d_k^2 g
        # (G_N.m:15-20), which is (d_i^2 + k0^2) g + delta: on the self voxel
        # it adds the overlap int T_l B_lp dV (h^3; h^3/12 for l = lp > 0)
        if i == j and tuple(m) == (0, 0, 0) and l == lp:
            val += h ** 3 * (1.0 if l == 0 else 1.0 / 12.0)
        return val
    # K: int T B d_j g = - T B n'_j g (basis faces) + T d'_j B g
    j = comp
    val = 0.0j
    for sp in (1, -1):
        kinds = [("P", sp) if d == j else ("V",) for d in range(3)]
        val -= _integral(boxes, ax(kinds, td, bd))
    if ba == j:
        val += _integral(boxes, ax([("V",)] * 3, td, zero)) / h
    return val


def near_field(res, k0, op, pwl, offsets, q_reg=8, q_sing=12):
    """Near entries {m: array (pwx, ncomp)} for the given offsets."""
    boxes = _Boxes(res, k0, q_reg, q_sing)
    pairs = PAIRS if pwl else PAIRS[:1]
    ncomp = 6 if op == "N" else 3
    out = {}
    for m in offsets:
        a = np.zeros((len(pairs), ncomp), dtype=np.complex128)
        for p, (l, lp) in enumerate(pairs):
            for c in range(ncomp):
                a[p, c] = near_entry(boxes, m, l, lp, op, c)
        out[tuple(m)] = a
    return out


def medium_rule(op: str, dims, medium_nq: int = 8):
    """Order and per-axis extent of the medium-range rule.  The order shrinks
    to min(n) on grids with some n <= 5 for N (assembly_N.m:36-38) but only
    with some n <= 2 for K (assembly_K.m:36-38); the rule covers the offsets
    [0, order) along each axis, clipped to the grid.  (Where the reference's
    order exceeds an axis length, its loop assigns past the n1 x n2 x n3 array
    and MATLAB grows it; inside the grid its values are these.)"""
    lim = 5 if op == "N" else 2
    nm = min(dims) if min(dims) <= lim else medium_nq
    return nm, tuple(min(nm, n) for n in dims)


def near_offsets(dims):
    """The offsets of the surface-surface (near) rule: {0,1} along every axis
    with n >= 2, {0} along the others (assembly_N.m:59-79, assembly_K.m:59-79)."""
    return list(itertools.product(*[(0, 1) if n >= 2 else (0,) for n in dims]))


def assemble(opr, op: str, k0: float, far_nq: int = 4, medium_nq: int = 8, near: bool = True):
    """Real-space Green tensor A[n1,n2,n3,pwx,ncomp] for VieOperator `opr`.

    Regimes as assembly_N.m:16-81 / assembly_K.m:17-80: far rule over every
    offset, medium rule over the first offsets (medium_rule), near offsets
    (near_offsets)."""
    n1, n2, n3 = opr.dims
    pwx = 10 if opr.pwl else 1
    ncomp = 6 if op == "N" else 3
    A = np.zeros((n1, n2, n3, pwx, ncomp), dtype=np.complex128, order="F")
    code = L.VIE_OP_N if op == "N" else L.VIE_OP_K

    def run(nq, e):
        x, w = np.polynomial.legendre.leggauss(nq)
        x = np.ascontiguousarray(x)
        w = np.ascontiguousarray(w)
        opr._check(opr.lib.vie_green_quadrature(opr.h, code, float(k0), nq, ctypes.c_void_p(x.ctypes.data),
                                                ctypes.c_void_p(w.ctypes.data), e[0], e[1], e[2],
                                                ctypes.c_void_p(A.ctypes.data), L.VIE_MEM_HOST))

    run(far_nq, (n1, n2, n3))
    nm, ext = medium_rule(op, (n1, n2, n3), medium_nq)
    run(nm, ext)
    if near:
        nf = near_field(opr.res, k0, op, opr.pwl, near_offsets((n1, n2, n3)))
        for m, a in nf.items():
            A[m[0], m[1], m[2], :, :] = a
    return A


def save(path, A: np.ndarray, op: str, dims, res: float, k0: float, pwl: bool) -> None:
    """Write an assembled real-space Green tensor with what it was built for
    (the set-up is the expensive part of a new frequency or grid; the
    reference recomputes it per run, assembly_N.m / assembly_K.m)."""
    A = np.asarray(A, dtype=np.complex128)
    want = tuple(dims) + ((10 if pwl else 1), (6 if op == "N" else 3))
    if A.shape != want:
        raise ValueError(f"A has shape {A.shape}, expected {want}")
    np.savez(path, A=A, op=op, dims=np.asarray(dims, dtype=np.int64), res=float(res), k0=float(k0), pwl=bool(pwl))


def load(path, op: str, dims, res: float, k0: float, pwl: bool) -> np.ndarray:
    """Read a tensor written by save(), checking that it was built for this
    operator, grid, voxel size, wavenumber and basis (no pickled objects)."""
    with np.load(path, allow_pickle=False) as f:
        got = (str(f["op"]), tuple(int(v) for v in f["dims"]), float(f["res"]), float(f["k0"]), bool(f["pwl"]))
        want = (op, tuple(int(v) for v in dims), float(res), float(k0), bool(pwl))
        if got[0] != want[0] or got[1] != want[1] or got[4] != want[4] or \
                abs(got[2] - want[2]) > 1e-12 * abs(want[2]) or abs(got[3] - want[3]) > 1e-12 * abs(want[3]):
            raise ValueError(f"{path} holds (op, dims, res, k0, pwl) = {got}, not {want}")
        return np.asfortranarray(f["A"])
