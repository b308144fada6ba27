"""Near-field N- and K-operator entries by the reference's surface-surface formulation.

Test infrastructure (the oracle), not the product: only tests/ and the golden
generator tests/golden/make_near_field.py use it.  For the offsets m in
{0,1}^3 the reference does not integrate the 6-D Galerkin volume integral of
the N or K kernel directly; Gauss's theorem moves it onto the 6 x 6 face pairs
of the two voxels, as a sum of 4 reduced kernels with constant coefficients:

  surface_surface_kernels_Nop.m:1-290   N: the assembly over face pairs and kernels
  surface_surface_coeff_Nop.m:1-95      N: the coefficient tables I1_co .. I4_co
  coefficients_Nop.m:1-107              N: one coefficient
  kernels_Nop.m:1-122                   N: the reduced kernels (non-singular pairs)
  surface_surface_kernels_Kop.m:1-267   K: the same four steps (DIRECTFN kernel types 5..8)
  surface_surface_coeff_Kop.m:1-82, coefficients_Kop.m:1-89, kernels_Kop.m:1-114
  points_const_4D.m:1-133               the 4-D Gauss points on a square pair
  singular_{ST,EA,VA}_lin.m             DIRECTFN for the touching pairs
  points_mapping.m:1-327                adjacency and corner order of each pair
  load_inputs.m:80                      Np_1D_near_S = 15

Each function below restates one of them (paths relative to
src/src_integral_equations/src_vie/src_operators_vie/).  The singular
integrals call the reference's own DIRECTFN C++ (Cpp_Assembly/src/src_cpp),
compiled from the reference checkout by oracle/Makefile into
oracle/_ref/libdirectfn.so; that library exists only where the reference
does (this is why the values are written once to tests/golden/).

points_mapping.m tabulates, for each of the eight m, the adjacency type of
every face pair and its corners in DIRECTFN's order.  face_pair() derives both
from the squares' shared corners instead: identical squares are ST, a shared
edge EA, a shared corner VA, none non-singular; the corners follow DIRECTFN's
convention (Geometry::EA/VA in classes.h: source square (r3, r4, r1, r2) and
observation square (r4, r3, r6, r5) around the shared edge r3-r4; source
(r3, r4, r1, r2) and observation (r3, r5, r6, r7) around the shared corner
r3).  The surface parametrisation uses the absolute Jacobian, so any corner
order with that structure gives the same integral; the adjacency tables are
checked against points_mapping.m's in tests/test_near_field.py.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libdirectfn.so")

#: outward normals of the six faces [-x, +x, -y, +y, -z, +z] (surface_surface_kernels_Nop.m:28-30)
FACE_N = np.array([[-1, 0, 0], [1, 0, 0], [0, -1, 0], [0, 1, 0], [0, 0, -1], [0, 0, 1]], dtype=float)
#: the six unique tensor components (p, q), 1-based (surface_surface_coeff_Nop.m:27-28)
P_IDX = (1, 1, 1, 2, 2, 3)
Q_IDX = (1, 2, 3, 2, 3, 3)
#: (l, lp) of the ten unique PWL pairs, 1-based (surface_surface_kernels_Nop.m:275-289)
PAIRS_L = ((1, 1), (2, 2), (3, 3), (4, 4), (1, 2), (1, 3), (1, 4), (2, 3), (2, 4), (3, 4))

ST, EA, VA = 1, 2, 3


def _lib():
    if not os.path.exists(LIB_PATH):
        raise FileNotFoundError(f"{LIB_PATH} is not built (make -C oracle, needs the reference checkout)")
    lib = ctypes.CDLL(LIB_PATH)
    P = ctypes.c_void_p
    for f in (lib.dfn_st, lib.dfn_ea, lib.dfn_va):
        f.argtypes = [P, ctypes.c_int, ctypes.c_double, ctypes.c_double, P, P, P, P, ctypes.c_int, ctypes.c_int,
                      ctypes.c_int, P]
        f.restype = ctypes.c_int
    return lib


def coefficient(dx, n, nq, p, q, l, lp, ker):
    """coefficients_Nop.m:1-107 (p, q, l, lp 1-based; ker 1..4)."""
    pp = np.zeros(3); pp[p - 1] = 1.0
    qq = np.zeros(3); qq[q - 1] = 1.0
    hv = lambda k: np.zeros(3) if k <= 1 else np.eye(3)[k - 2] / dx
    if ker == 1:
        return float(np.dot(np.cross(n, pp), np.cross(nq, qq)))
    if ker == 2:
        return float(np.dot(np.cross(np.cross(hv(l), pp), qq), n))
    if ker == 3:
        return float(np.dot(np.cross(pp, np.cross(hv(lp), qq)), n))
    a = np.cross(hv(l), pp)
    e = np.eye(3)
    return float(sum(np.dot(np.cross(a, e[i]), n) * np.dot(hv(lp) * qq[i], nq) for i in range(3)))


def coefficient_tables(dx, L):
    """surface_surface_coeff_Nop.m: I1 (6,6,6), I2 (6,6,6,L), I3 (6,6,6,L), I4 (6,6,6,L,L)."""
    I1 = np.zeros((6, 6, 6)); I2 = np.zeros((6, 6, 6, L)); I3 = np.zeros((6, 6, 6, L)); I4 = np.zeros((6, 6, 6, L, L))
    for f in range(6):
        for fp in range(6):
            n, nq = FACE_N[f], FACE_N[fp]
            for i in range(6):
                p, q = P_IDX[i], Q_IDX[i]
                I1[f, fp, i] = coefficient(dx, n, nq, p, q, 0, 0, 1)
                for l in range(1, L + 1):
                    I2[f, fp, i, l - 1] = coefficient(dx, n, nq, p, q, l, 0, 2)
                    I3[f, fp, i, l - 1] = coefficient(dx, n, nq, p, q, 0, l, 3)
                    for lp in range(1, L + 1):
                        I4[f, fp, i, l - 1, lp - 1] = coefficient(dx, n, nq, p, q, l, lp, 4)
    return I1, I2, I3, I4


def square_points(A, B, C, D, r_ms, r_ns, dx, n, nq):
    """points_const_4D.m: observation square through r_ms with normal n (its
    two free coordinates, in x, y, z order, take A and B) and source square
    through r_ns with normal nq (C and D)."""
    plane = int(np.nonzero(n)[0][0])
    plane_p = int(np.nonzero(nq)[0][0])
    obs = [None] * 3
    src = [None] * 3
    obs[plane] = np.full(A.shape, r_ms[plane])
    src[plane_p] = np.full(A.shape, r_ns[plane_p])
    free = [d for d in range(3) if d != plane]
    free_p = [d for d in range(3) if d != plane_p]
    obs[free[0]] = r_ms[free[0]] + dx / 2 * A
    obs[free[1]] = r_ms[free[1]] + dx / 2 * B
    src[free_p[0]] = r_ns[free_p[0]] + dx / 2 * C
    src[free_p[1]] = r_ns[free_p[1]] + dx / 2 * D
    return obs[0], obs[1], obs[2], src[0], src[1], src[2]


def reduced_kernel(X, Y, Z, Xp, Yp, Zp, k0, dx, r_m, r_n, l, lp, nq, ker):
    """kernels_Nop.m:1-122: the reduced surface-surface kernel ker at the points
    (r_m, r_n: voxel centres; l, lp 1-based monomials, 0 where unused)."""
    mono = lambda k, P3, c: 1.0 if k <= 1 else (P3[k - 2] - c[k - 2]) / dx
    R = np.sqrt((X - Xp) ** 2 + (Y - Yp) ** 2 + (Z - Zp) ** 2)
    G = np.exp(-1j * k0 * R) / (4 * np.pi * R)
    G0 = 1.0 / (4 * np.pi * R)
    if ker == 1:
        return mono(l, (X, Y, Z), r_m) * mono(lp, (Xp, Yp, Zp), r_n) * G
    if ker in (2, 3):
        g = (-1j * k0 * G / R - G / R ** 2 + G0 / R ** 2) / (1j * k0) ** 2
        q2 = nq[0] * (X - Xp) * g + nq[1] * (Y - Yp) * g + nq[2] * (Z - Zp) * g
        return (mono(lp, (Xp, Yp, Zp), r_n) if ker == 2 else mono(l, (X, Y, Z), r_m)) * q2
    return (G - G0) / (1j * k0) ** 2


def _corners(centre, face, dx):
    """Corners of face `face` of the voxel at `centre`, in cyclic order."""
    n = FACE_N[face]
    d = int(np.nonzero(n)[0][0])
    u, v = [k for k in range(3) if k != d]
    out = []
    for a, b in ((-1, -1), (1, -1), (1, 1), (-1, 1)):
        c = np.array(centre, dtype=float) + n * dx / 2
        c[u] += a * dx / 2
        c[v] += b * dx / 2
        out.append(c)
    return out


def _rotate_to(cyc, first, second):
    """Cyclic order of the square `cyc` starting first -> second."""
    i = next(k for k, c in enumerate(cyc) if np.allclose(c, first))
    if np.allclose(cyc[(i + 1) % 4], second):
        return [cyc[(i + k) % 4] for k in range(4)]
    return [cyc[(i - k) % 4] for k in range(4)]


def face_pair(m, dx, face, face_p):
    """Adjacency type (0 none, ST, EA, VA) of observation face `face` of the
    voxel at (m - 1) dx and source face `face_p` of the voxel at 0, and the
    corners r1 .. r7 in DIRECTFN's order (the role of points_mapping.m)."""
    r_m = (np.asarray(m) - 1) * dx
    obs = _corners(r_m, face, dx)
    src = _corners(np.zeros(3), face_p, dx)
    shared = [c for c in obs if any(np.allclose(c, s, atol=1e-12 * dx) for s in src)]
    pts = np.zeros((7, 3))
    if len(shared) == 4:
        pts[:4] = np.array(src)
        return ST, pts
    if len(shared) == 2:
        a, b = shared
        q = _rotate_to(src, a, b)           # (r3, r4, r1, r2): r3 = a, r4 = b
        p = _rotate_to(obs, b, a)           # (r4, r3, r6, r5)
        pts[2], pts[3], pts[0], pts[1] = q
        pts[5], pts[4] = p[2], p[3]
        return EA, pts
    if len(shared) == 1:
        a = shared[0]
        q = _rotate_to(src, a, src[(next(k for k, c in enumerate(src) if np.allclose(c, a)) + 1) % 4])
        p = _rotate_to(obs, a, obs[(next(k for k, c in enumerate(obs) if np.allclose(c, a)) + 1) % 4])
        pts[2], pts[3], pts[0], pts[1] = q  # (r3, r4, r1, r2)
        pts[2], pts[4], pts[5], pts[6] = p  # (r3, r5, r6, r7)
        return VA, pts
    return 0, pts


def _singular(lib, typ, pts, Np, k0, dx, n, nq, ker, l, lp):
    """singular_{ST,EA,VA}_lin.m: DIRECTFN with the square centres of the
    MATLAB wrappers; l, lp 0-based as the C++ takes them."""
    if typ == ST:
        rq_c = (pts[0] + pts[2]) / 2
        rp_c = rq_c
    elif typ == EA:
        rq_c = (pts[0] + pts[2]) / 2
        rp_c = (pts[3] + pts[5]) / 2
    else:
        rq_c = (pts[0] + pts[2]) / 2
        rp_c = (pts[2] + pts[5]) / 2
    arrs = [np.ascontiguousarray(a, dtype=float) for a in (pts, rq_c, rp_c, nq, n)]
    out = np.zeros(2)
    fn = {ST: lib.dfn_st, EA: lib.dfn_ea, VA: lib.dfn_va}[typ]
    ptr = lambda a: ctypes.c_void_p(a.ctypes.data)
    # the MATLAB wrappers pass (nq, np) = (source normal, observation normal)
    fn(ptr(arrs[0]), int(Np), float(k0), float(dx), ptr(arrs[1]), ptr(arrs[2]), ptr(arrs[3]), ptr(arrs[4]),
       int(ker), int(l), int(lp), ptr(out))
    return complex(out[0], out[1])


def near_N(m, dx, k0, L, Np=15, lib=None):
    """surface_surface_kernels_Nop.m:1-290 at voxel offset m (1-based as in the
    reference, m in {1,2}^3): the (pwx, 6) near interactions, pwx = 1 (L = 1)
    or 10 (L = 4), rows in the pair order of PAIRS_L."""
    lib = lib or _lib()
    I1c, I2c, I3c, I4c = coefficient_tables(dx, L)
    x, w = np.polynomial.legendre.leggauss(Np)
    Ag, Bg, Cg, Dg = np.meshgrid(x, x, x, x, indexing="ij")
    W = np.einsum("a,b,c,d->abcd", w, w, w, w).reshape(-1)
    A_, B_, C_, D_ = (v.reshape(-1) for v in (Ag, Bg, Cg, Dg))
    J = (dx / 2) ** 4
    r_m = (np.asarray(m, dtype=float) - 1) * dx
    r_n = np.zeros(3)
    ISS = np.zeros((L, L, 6), dtype=np.complex128)
    for f in range(6):
        for fp in range(6):
            n, nq = FACE_N[f], FACE_N[fp]
            typ, pts = face_pair(m, dx, f, fp)
            if typ == 0:
                Xs = square_points(A_, B_, C_, D_, r_m + n * dx / 2, r_n + nq * dx / 2, dx, n, nq)

                def val(l, lp, ker):
                    return np.sum(W * reduced_kernel(*Xs, k0, dx, r_m, r_n, l, lp, nq, ker)) * J
            else:
                def val(l, lp, ker):
                    return _singular(lib, typ, pts, Np, k0, dx, n, nq, ker, max(l - 1, 0), max(lp - 1, 0))
            # kernel 1: T B G, every (l, lp)
            for l in range(1, L + 1):
                for lp in range(1, L + 1):
                    if np.any(I1c[f, fp] != 0):
                        v = val(l, lp, 1)
                        ISS[l - 1, lp - 1] += I1c[f, fp] * v
            # kernel 2: basis monomial lp, the test's gradient in the coefficient
            for lp in range(1, L + 1):
                if np.any(I2c[f, fp] != 0):
                    v = val(0, lp, 2)
                    for l in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I2c[f, fp, :, l - 1] * v
            # kernel 3: test monomial l
            for l in range(1, L + 1):
                if np.any(I3c[f, fp] != 0):
                    v = val(l, 0, 3)
                    for lp in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I3c[f, fp, :, lp - 1] * v
            # kernel 4: (G - G0) / (j k0)^2
            if np.any(I4c[f, fp] != 0):
                v = val(0, 0, 4)
                for l in range(1, L + 1):
                    for lp in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I4c[f, fp, :, l - 1, lp - 1] * v
    if L == 1:
        return ISS[0, 0][None, :]
    return np.stack([ISS[l - 1, lp - 1] for l, lp in PAIRS_L])


# ---------------------------------------------------------------------------
# K operator (surface_surface_kernels_Kop.m, surface_surface_coeff_Kop.m,
# coefficients_Kop.m, kernels_Kop.m; assembly_K.m:54-80)
# ---------------------------------------------------------------------------

#: the three unique (p, q) of the antisymmetric K tensor, 1-based (surface_surface_coeff_Kop.m:10-11)
P_IDX_K = (3, 1, 2)
Q_IDX_K = (2, 3, 1)


def coefficient_K(dx, k0, n, nq, p, q, l, lp, ker):
    """coefficients_Kop.m:1-89 (n observation, nq source normal; p, q, l, lp
    1-based, 0 where unused; ker 1..4)."""
    pp = np.zeros(3); pp[p - 1] = 1.0
    qq = np.zeros(3); qq[q - 1] = 1.0
    hv = lambda k: np.zeros(3) if k <= 1 else np.eye(3)[k - 2] / dx
    if ker == 1:
        return complex(np.dot(np.cross(pp, qq), nq))
    if ker == 2:
        return complex(np.dot(nq, np.dot(np.cross(qq, pp), hv(lp)) * hv(l)) / (1j * k0) ** 2)
    if ker == 3:
        return complex(np.dot(n, nq) * np.dot(hv(lp), np.cross(pp, qq)))
    e = np.eye(3)
    return complex(sum(np.dot(np.cross(qq, e[i]), nq) * np.dot(hv(l) * pp[i], n) for i in range(3)))


def coefficient_tables_K(dx, k0, L):
    """surface_surface_coeff_Kop.m: I1 (6,6,3), I2 (6,6,3,L,L), I3 (6,6,3,L) over
    lp, I4 (6,6,3,L) over l."""
    I1 = np.zeros((6, 6, 3), complex); I2 = np.zeros((6, 6, 3, L, L), complex)
    I3 = np.zeros((6, 6, 3, L), complex); I4 = np.zeros((6, 6, 3, L), complex)
    for f in range(6):
        for fp in range(6):
            n, nq = FACE_N[f], FACE_N[fp]
            for i in range(3):
                p, q = P_IDX_K[i], Q_IDX_K[i]
                I1[f, fp, i] = coefficient_K(dx, k0, n, nq, p, q, 0, 0, 1)
                for l in range(1, L + 1):
                    I3[f, fp, i, l - 1] = coefficient_K(dx, k0, n, nq, p, q, 0, l, 3)
                    I4[f, fp, i, l - 1] = coefficient_K(dx, k0, n, nq, p, q, l, 0, 4)
                    for lp in range(1, L + 1):
                        I2[f, fp, i, l - 1, lp - 1] = coefficient_K(dx, k0, n, nq, p, q, l, lp, 2)
    return I1, I2, I3, I4


def reduced_kernel_K(X, Y, Z, Xp, Yp, Zp, k0, dx, r_m, r_n, l, lp, n, ker):
    """kernels_Kop.m:1-114 (n: the OBSERVATION face normal, as
    surface_surface_kernels_Kop.m passes it; l, lp 1-based, 0 where unused)."""
    mono = lambda k, P3, c: 1.0 if k <= 1 else (P3[k - 2] - c[k - 2]) / dx
    R = np.sqrt((X - Xp) ** 2 + (Y - Yp) ** 2 + (Z - Zp) ** 2)
    G = np.exp(-1j * k0 * R) / (4 * np.pi * R)
    G0 = 1.0 / (4 * np.pi * R)
    if ker in (1, 2):
        g = (-1j * k0 * G / R - G / R ** 2 + G0 / R ** 2) / (1j * k0) ** 2
        d = (X - Xp, Y - Yp, Z - Zp)
        if ker == 1:
            q2 = sum(n[k] * d[k] * g for k in range(3))
            return mono(l, (X, Y, Z), r_m) * mono(lp, (Xp, Yp, Zp), r_n) * q2
        return sum(n[k] * (d[k] * g - d[k] / 2 * G0) for k in range(3))
    if ker == 3:
        return mono(l, (X, Y, Z), r_m) * (G - G0) / (1j * k0) ** 2
    return mono(lp, (Xp, Yp, Zp), r_n) * (G - G0) / (1j * k0) ** 2


def near_K(m, dx, k0, L, Np=15, lib=None):
    """surface_surface_kernels_Kop.m:1-267 at voxel offset m (1-based, m in
    {1,2}^3): the (pwx, 3) near interactions, columns the (p, q) of
    P_IDX_K / Q_IDX_K, rows in the pair order of PAIRS_L.  The singular
    pairs take DIRECTFN kernel types 5..8."""
    lib = lib or _lib()
    I1c, I2c, I3c, I4c = coefficient_tables_K(dx, k0, L)
    x, w = np.polynomial.legendre.leggauss(Np)
    Ag, Bg, Cg, Dg = np.meshgrid(x, x, x, x, indexing="ij")
    W = np.einsum("a,b,c,d->abcd", w, w, w, w).reshape(-1)
    A_, B_, C_, D_ = (v.reshape(-1) for v in (Ag, Bg, Cg, Dg))
    J = (dx / 2) ** 4
    r_m = (np.asarray(m, dtype=float) - 1) * dx
    r_n = np.zeros(3)
    ISS = np.zeros((L, L, 3), dtype=np.complex128)
    for f in range(6):
        for fp in range(6):
            n, nq = FACE_N[f], FACE_N[fp]
            typ, pts = face_pair(m, dx, f, fp)
            if typ == 0:
                Xs = square_points(A_, B_, C_, D_, r_m + n * dx / 2, r_n + nq * dx / 2, dx, n, nq)

                def val(l, lp, ker):
                    return np.sum(W * reduced_kernel_K(*Xs, k0, dx, r_m, r_n, l, lp, n, ker)) * J
            else:
                def val(l, lp, ker):
                    return _singular(lib, typ, pts, Np, k0, dx, n, nq, ker + 4, max(l - 1, 0), max(lp - 1, 0))
            # kernel 1: T B (n . F), every (l, lp)
            if np.any(I1c[f, fp] != 0):
                for l in range(1, L + 1):
                    for lp in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I1c[f, fp] * val(l, lp, 1)
            # kernel 2: n . (F - R G0 / 2), both gradients in the coefficient
            if np.any(I2c[f, fp] != 0):
                v = val(0, 0, 2)
                for l in range(1, L + 1):
                    for lp in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I2c[f, fp, :, l - 1, lp - 1] * v
            # kernel 3: test monomial l, the basis gradient (lp) in the coefficient
            if np.any(I3c[f, fp] != 0):
                for l in range(1, L + 1):
                    v = val(l, 0, 3)
                    for lp in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I3c[f, fp, :, lp - 1] * v
            # kernel 4: basis monomial lp, the test gradient (l) in the coefficient
            if np.any(I4c[f, fp] != 0):
                for lp in range(1, L + 1):
                    v = val(0, lp, 4)
                    for l in range(1, L + 1):
                        ISS[l - 1, lp - 1] += I4c[f, fp, :, l - 1] * v
    if L == 1:
        return ISS[0, 0][None, :]
    return np.stack([ISS[l - 1, lp - 1] for l, lp in PAIRS_L])
