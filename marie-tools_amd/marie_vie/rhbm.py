"""RHBM body models: the on-disk struct MARIE reads, and its downsampling.

An RHBM file is a MAT v5 file holding one struct `RHBM` with fields
  r          (n1, n2, n3, 3) voxel centres, x fastest
  epsilon_r  (n1, n2, n3) relative permittivity
  sigma_e    (n1, n2, n3) conductivity [S/m]
  rho        (n1, n2, n3) density
  idxS       (nS, 1) 1-based linear (column-major) indices of the body voxels
  name       text
The grid and the scatterer set follow geo_body_domain.m:5-19 (n_d = size(r, d),
res = |r(2,1,1,1) - r(1,1,1,1)|, mask(idxS) = true); downsample() follows
RHBM_downsample.m:1-141.  Only data is read (scipy.io.loadmat of MAT v5).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np
import scipy.io

from .phantoms import Phantom


@dataclass
class RHBM:
    r: np.ndarray
    epsilon_r: np.ndarray
    sigma_e: np.ndarray
    rho: np.ndarray
    idxS: np.ndarray       # 0-based, sorted
    name: str = ""

    @property
    def dims(self) -> tuple:
        return tuple(int(v) for v in self.r.shape[:3])

    @property
    def res(self) -> float:
        """abs(r(2,1,1,1) - r(1,1,1,1)) (geo_body_domain.m:12)."""
        return float(abs(self.r[1, 0, 0, 0] - self.r[0, 0, 0, 0]))

    def phantom(self) -> Phantom:
        """The voxel model the operator takes (materials outside idxS are
        air, as the contrast is applied on idxS only)."""
        mask = np.zeros(int(np.prod(self.dims)), dtype=bool)
        mask[self.idxS] = True
        mask = mask.reshape(self.dims, order="F")
        eps = np.where(mask, np.asarray(self.epsilon_r, dtype=np.float64), 1.0)
        sig = np.where(mask, np.asarray(self.sigma_e, dtype=np.float64), 0.0)
        origin = tuple(float(v) for v in self.r[0, 0, 0, :])
        return Phantom(self.dims, self.res, eps, sig, np.asarray(self.idxS, dtype=np.int64), origin)


def _field(s, name):
    v = s[name]
    while isinstance(v, np.ndarray) and v.dtype == object and v.size == 1:
        v = v.item()
    return v


def load(path: str) -> RHBM:
    """Read the RHBM struct (geo_body_domain.m:5-6 `load(rhbm)`)."""
    d = scipy.io.loadmat(path)
    if "RHBM" not in d:
        raise ValueError(f"{path}: no RHBM struct")
    s = d["RHBM"][0, 0]
    r = np.asarray(_field(s, "r"), dtype=np.float64)
    if r.ndim != 4 or r.shape[3] != 3:
        raise ValueError("RHBM.r must be n1 x n2 x n3 x 3")
    idx = np.asarray(_field(s, "idxS")).reshape(-1).astype(np.int64) - 1
    if idx.size and (idx.min() < 0 or idx.max() >= np.prod(r.shape[:3])):
        raise ValueError("RHBM.idxS out of range")
    name = ""
    if "name" in s.dtype.names:
        nm = _field(s, "name")
        name = str(nm[0]) if isinstance(nm, np.ndarray) and nm.size else str(nm)
    return RHBM(r, np.asarray(_field(s, "epsilon_r")), np.asarray(_field(s, "sigma_e")),
                np.asarray(_field(s, "rho")), np.sort(idx), name)


def save(path: str, m: RHBM) -> None:
    """Write the struct MARIE loads (idxS 1-based, column)."""
    st = {"r": m.r, "epsilon_r": m.epsilon_r, "sigma_e": m.sigma_e, "rho": m.rho,
          "idxS": (np.asarray(m.idxS, dtype=np.float64) + 1).reshape(-1, 1), "name": m.name}
    scipy.io.savemat(path, {"RHBM": st}, do_compression=True)


def _block(A, f, shape2, red):
    nx, ny, nz = shape2
    B = np.asarray(A).reshape((f, nx, f, ny, f, nz), order="F")
    return red(B, axis=(0, 2, 4))


def _pad(A, f):
    """Edge-replicate each axis up to a multiple of f (pad_to_multiple3d)."""
    pads = [((-n) % f) for n in A.shape]
    return np.pad(A, [(0, p) for p in pads], mode="edge")


def _cast(v, ref):
    """MATLAB cast(v, class(ref)): integer classes round half away from zero
    and saturate."""
    dt = np.asarray(ref).dtype
    if dt.kind == "f":
        return np.asarray(v).astype(dt)
    if dt.kind in "iu":
        info = np.iinfo(dt)
        r = np.sign(v) * np.floor(np.abs(v) + 0.5)
        return np.clip(r, info.min, info.max).astype(dt)
    return np.asarray(v).astype(dt)


def downsample(m: RHBM, factor: int, use_nearest: bool = False, mask_policy: str = "any",
               preserve_fov: str = "crop") -> RHBM:
    """RHBM_downsample.m: isotropic integer downsampling.

    Block mean of epsilon_r, sigma_e, rho (or every factor-th voxel with
    use_nearest); the body mask of a coarse voxel is any / all of its fine
    voxels; 'crop' drops trailing slices, 'pad' edge-replicates them; the
    coarse centres are r0 + k * factor * d along each axis (d the median fine
    spacing, RHBM_downsample.m:100-110)."""
    if factor < 1 or int(factor) != factor:
        raise ValueError("factor must be a positive integer")
    f = int(factor)
    dims = m.dims
    mask = np.zeros(int(np.prod(dims)), dtype=bool)
    mask[m.idxS] = True
    mask = mask.reshape(dims, order="F")
    A = [np.asarray(m.epsilon_r), np.asarray(m.sigma_e), np.asarray(m.rho), mask]
    if preserve_fov == "crop":
        ne = [(n // f) * f for n in dims]
        A = [a[:ne[0], :ne[1], :ne[2]] for a in A]
    elif preserve_fov == "pad":
        A = [_pad(a, f) for a in A]
    else:
        raise ValueError("preserve_fov must be 'crop' or 'pad'")
    n2 = tuple(s // f for s in A[0].shape)
    if use_nearest:
        eps2, sig2, rho2, m2 = [a[::f, ::f, ::f] for a in A]
    else:
        eps2, sig2, rho2 = [_block(a.astype(np.float64), f, n2, np.nanmean) for a in A[:3]]
        red = np.any if mask_policy == "any" else np.all
        if mask_policy not in ("any", "all"):
            raise ValueError("mask_policy must be 'any' or 'all'")
        m2 = _block(A[3], f, n2, red)
    d = [float(np.median(np.diff(m.r[:, 0, 0, 0]))), float(np.median(np.diff(m.r[0, :, 0, 1]))),
         float(np.median(np.diff(m.r[0, 0, :, 2])))]
    x0 = m.r[0, 0, 0, :]
    lines = [x0[k] + np.arange(n2[k]) * (f * d[k]) for k in range(3)]
    X, Y, Z = np.meshgrid(*lines, indexing="ij")
    r2 = np.stack([X, Y, Z], axis=-1).astype(m.r.dtype)
    return RHBM(r2, _cast(eps2, m.epsilon_r), _cast(sig2, m.sigma_e), _cast(rho2, m.rho),
                np.nonzero(m2.reshape(-1, order="F"))[0].astype(np.int64),
                f"{m.name}_ds{f}" if m.name else f"RHBM_ds{f}")
