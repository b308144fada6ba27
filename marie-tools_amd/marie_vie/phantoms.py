"""Synthetic voxel phantoms and excitations for the benchmark workloads.

Voxel centres are x_i = (i - (n+1)/2) * res (i = 1..n) on each axis and a
voxel belongs to a body when its centre does.  Random choices use
numpy.default_rng with the seed given.  Materials become Mr, Mc by
em_assembly.m:14-15 and the incident field is projected on the voxel basis
(constant = voxel average; linear moments divided by the Gram weight
res^3/12, mvp_G_pwl.m:5).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

C0 = 299792458.0
MU0 = 4e-7 * math.pi
E0 = 1.0 / (C0 ** 2 * MU0)


def constants(freq_hz: float) -> dict:
    """k0 and ce = j omega e0 (em_constants.m:127-139)."""
    omega = 2 * math.pi * freq_hz
    return {"k0": omega / C0, "ce": 1j * omega * E0, "omega": omega}


@dataclass
class Phantom:
    dims: tuple
    res: float
    eps_r: np.ndarray     # (n1, n2, n3)
    sigma: np.ndarray     # (n1, n2, n3)
    idxS: np.ndarray      # 0-based Fortran-order scatterer voxels
    origin: tuple | None = None   # centre of voxel (0, 0, 0); None: the grid is centred at 0

    def centres(self):
        if self.origin is not None:
            return [o + np.arange(n) * self.res for o, n in zip(self.origin, self.dims)]
        return [(np.arange(1, n + 1) - (n + 1) / 2) * self.res for n in self.dims]

    def material(self, freq_hz: float):
        ce = constants(freq_hz)["ce"]
        Mr = self.eps_r + self.sigma / ce
        return Mr, Mr - 1.0


def _grid(dims, res):
    xs = [(np.arange(1, n + 1) - (n + 1) / 2) * res for n in dims]
    return np.meshgrid(*xs, indexing="ij")


def _finish(dims, res, eps, sig, mask):
    idxS = np.nonzero(mask.reshape(-1, order="F"))[0].astype(np.int64)
    eps = np.where(mask, eps, 1.0)
    sig = np.where(mask, sig, 0.0)
    return Phantom(tuple(dims), res, eps, sig, idxS)


def sphere(dims=(32, 32, 32), res=0.005, radius=0.07, eps_r=50.0, sigma=0.6) -> Phantom:
    """Homogeneous dielectric sphere at the origin (32^3 grid of 5 mm voxels)."""
    X, Y, Z = _grid(dims, res)
    mask = X ** 2 + Y ** 2 + Z ** 2 <= radius ** 2
    return _finish(dims, res, np.full(dims, eps_r), np.full(dims, sigma), mask)


def ellipsoid_regions(dims=(64, 64, 80), res=0.004, semi=(0.080, 0.095, 0.110), ncell=12,
                      eps_range=(40.0, 80.0), sig_range=(0.3, 0.9), seed=1002) -> Phantom:
    """Ellipsoid head phantom split into seeded piecewise-constant (Voronoi) regions."""
    rng = np.random.default_rng(seed)
    X, Y, Z = _grid(dims, res)
    mask = (X / semi[0]) ** 2 + (Y / semi[1]) ** 2 + (Z / semi[2]) ** 2 <= 1.0
    # region seeds: uniform points inside the ellipsoid (rejection sampling)
    seeds = []
    while len(seeds) < ncell:
        p = (rng.random(3) * 2 - 1) * np.array(semi)
        if ((p / np.array(semi)) ** 2).sum() <= 1.0:
            seeds.append(p)
    seeds = np.array(seeds)
    eps_c = rng.uniform(*eps_range, size=ncell)
    sig_c = rng.uniform(*sig_range, size=ncell)
    P = np.stack([X, Y, Z], axis=-1)
    d2 = ((P[..., None, :] - seeds) ** 2).sum(-1)
    lab = d2.argmin(-1)
    return _finish(dims, res, eps_c[lab], sig_c[lab], mask)


def layered_torso(dims=(96, 96, 112), res=0.003, semi=(0.160, 0.110, 0.280), seed=1004) -> Phantom:
    """Torso ellipsoid (clipped by the grid box) with seeded fat / muscle / organ shells."""
    rng = np.random.default_rng(seed)
    X, Y, Z = _grid(dims, res)
    r = np.sqrt((X / semi[0]) ** 2 + (Y / semi[1]) ** 2 + (Z / semi[2]) ** 2)
    mask = r <= 1.0
    fat = (rng.uniform(5, 12), rng.uniform(0.04, 0.08))
    muscle = (rng.uniform(50, 60), rng.uniform(0.6, 0.8))
    organ = (rng.uniform(55, 70), rng.uniform(0.5, 0.9))
    t_fat = rng.uniform(0.05, 0.10)
    t_mus = rng.uniform(0.15, 0.25)
    eps = np.where(r > 1 - t_fat, fat[0], np.where(r > 1 - t_fat - t_mus, muscle[0], organ[0]))
    sig = np.where(r > 1 - t_fat, fat[1], np.where(r > 1 - t_fat - t_mus, muscle[1], organ[1]))
    return _finish(dims, res, eps, sig, mask)


# full body: (centre, semi-axes) of the ellipsoids, metres (legs, torso, head)
BODY_PARTS = (((-0.09, 0.0, -0.19), (0.07, 0.07, 0.13)),
              ((0.09, 0.0, -0.19), (0.07, 0.07, 0.13)),
              ((0.0, 0.0, 0.05), (0.17, 0.11, 0.15)),
              ((0.0, 0.0, 0.24), (0.075, 0.09, 0.08)))


def body_stack(dims=(128, 128, 160), res=0.004, ncell=32, eps_range=(5.0, 75.0), sig_range=(0.05, 0.9),
               seed=1005) -> Phantom:
    """Full-body ellipsoid stack (two legs, torso, head; 512 x 512 x 640 mm at
    4 mm) with seeded piecewise-constant tissue: ncell Voronoi regions with
    seeds inside the body, eps_r and sigma uniform in the given ranges.  The
    regions are labelled one z slab at a time (bounded memory)."""
    rng = np.random.default_rng(seed)
    xs = [(np.arange(1, n + 1) - (n + 1) / 2) * res for n in dims]

    def inside(px, py, pz):
        m = np.zeros(np.broadcast(px, py, pz).shape, bool)
        for c, a in BODY_PARTS:
            m |= ((px - c[0]) / a[0]) ** 2 + ((py - c[1]) / a[1]) ** 2 + ((pz - c[2]) / a[2]) ** 2 <= 1.0
        return m

    lo = np.array([x[0] for x in xs])
    hi = np.array([x[-1] for x in xs])
    seeds = []
    while len(seeds) < ncell:
        p = lo + rng.random(3) * (hi - lo)
        if inside(*p):
            seeds.append(p)
    seeds = np.array(seeds)
    eps_c = rng.uniform(*eps_range, size=ncell)
    sig_c = rng.uniform(*sig_range, size=ncell)
    X, Y = np.meshgrid(xs[0], xs[1], indexing="ij")
    eps = np.empty(dims)
    sig = np.empty(dims)
    mask = np.empty(dims, bool)
    for k, z in enumerate(xs[2]):
        mask[:, :, k] = inside(X, Y, z)
        d2 = (X[..., None] - seeds[:, 0]) ** 2 + (Y[..., None] - seeds[:, 1]) ** 2 + (z - seeds[:, 2]) ** 2
        lab = d2.argmin(-1)
        eps[:, :, k] = eps_c[lab]
        sig[:, :, k] = sig_c[lab]
    return _finish(dims, res, eps, sig, mask)


# ---------------------------------------------------------------------------
# excitations: field coefficients on the voxel basis, packed like idx_solution
# ---------------------------------------------------------------------------
def project(phantom: Phantom, efield, pwl: bool, nq: int = 3) -> np.ndarray:
    """Project a field E(x, y, z) -> (Ex, Ey, Ez) on the PWC/PWL voxel basis at
    the scatterer voxels; returns the packed coefficient vector (ql * nS)."""
    x, w = np.polynomial.legendre.leggauss(nq)
    h = phantom.res
    cx, cy, cz = phantom.centres()
    n1, n2 = phantom.dims[0], phantom.dims[1]
    s = phantom.idxS
    i1 = s % n1
    i2 = (s // n1) % n2
    i3 = s // (n1 * n2)
    X0, Y0, Z0 = cx[i1], cy[i2], cz[i3]
    ls = 4 if pwl else 1
    coef = np.zeros((3, ls, s.size), dtype=np.complex128)
    for a, wa in zip(x, w):
        for b, wb in zip(x, w):
            for c, wc in zip(x, w):
                ww = wa * wb * wc / 8.0              # average over the cube
                E = efield(X0 + a * h / 2, Y0 + b * h / 2, Z0 + c * h / 2)
                mono = [1.0, a / 2, b / 2, c / 2]
                for p in range(3):
                    for l in range(ls):
                        coef[p, l] += ww * mono[l] * E[p]
    if pwl:
        coef[:, 1:, :] *= 12.0                     # divide by the Gram weight 1/12 of xi^2
    return coef.reshape(3 * ls, s.size).reshape(-1)


def plane_wave(k0: float):
    """x-polarised plane wave travelling along +z: E = x_hat exp(-i k0 z)."""
    def f(x, y, z):
        e = np.exp(-1j * k0 * z)
        return (e, np.zeros_like(e), np.zeros_like(e))
    return f


def surface_patterns(k0: float, n: int, radius: float, zspan: float, nelem: int, seed: int):
    """n random-phase surface-current patterns on a cylinder (radius, |z| <=
    zspan): each is the field of nelem seeded small tangential current
    elements (unit moment) with seeded phases exp(i phi).  Returns field
    callables like dipoles()."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        elems = dipoles(k0, nelem, radius, zspan, int(rng.integers(1 << 31)))
        ph = np.exp(1j * rng.uniform(0, 2 * math.pi, size=nelem))

        def f(x, y, z, elems=elems, ph=ph):
            ex = ey = ez = 0.0
            for e, c in zip(elems, ph):
                a, b, d = e(x, y, z)
                ex = ex + c * a
                ey = ey + c * b
                ez = ez + c * d
            return ex, ey, ez
        out.append(f)
    return out


def dipoles(k0: float, n: int, radius: float, zspan: float, seed: int):
    """n Hertzian dipoles on a cylinder (radius, |z| <= zspan), tangential
    orientation, seeded angles; returns a list of field callables (closed form
    near + far field of a unit dipole moment)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        phi = rng.uniform(0, 2 * math.pi)
        z0 = rng.uniform(-zspan, zspan)
        pos = np.array([radius * math.cos(phi), radius * math.sin(phi), z0])
        ang = rng.uniform(0, 2 * math.pi)
        t_phi = np.array([-math.sin(phi), math.cos(phi), 0.0])
        t_z = np.array([0.0, 0.0, 1.0])
        pdir = math.cos(ang) * t_phi + math.sin(ang) * t_z

        def f(x, y, z, pos=pos, pdir=pdir):
            rx, ry, rz = x - pos[0], y - pos[1], z - pos[2]
            R = np.sqrt(rx ** 2 + ry ** 2 + rz ** 2)
            ux, uy, uz = rx / R, ry / R, rz / R
            kr = k0 * R
            g = np.exp(-1j * kr) / (4 * math.pi * R)
            # exp(+i omega t) convention: E ~ g [ (k^2 - 1/R^2 - ik/R) p
            #                                      + (-k^2 + 3/R^2 + 3ik/R)(p.u) u ]
            a = k0 ** 2 - 1.0 / R ** 2 - 1j * k0 / R
            b = -k0 ** 2 + 3.0 / R ** 2 + 3j * k0 / R
            pu = pdir[0] * ux + pdir[1] * uy + pdir[2] * uz
            return (g * (a * pdir[0] + b * pu * ux), g * (a * pdir[1] + b * pu * uy),
                    g * (a * pdir[2] + b * pu * uz))
        out.append(f)
    return out
