"""Field recovery from the solved polarisation currents (host side of the
library's N and K operators).

  E = e_inc + mask . invG( (N J - G J) / ce )   (em_efield_vie_excitation.m:18-24)
  B = b_inc + mask . invG( K J )                (em_hfield_vie_excitation.m:16-23)

J, e_inc, b_inc are packed vectors in idx_solution order; the operators run
on the device through VieOperator (dense [n1,n2,n3,ql] fields).  invG(G J)
is J, so the E update is (invG N J - J) / ce.
"""
from __future__ import annotations

import numpy as np

from .operator import VieOperator


def _gram(op: VieOperator) -> np.ndarray:
    h = op.res
    if op.pwl:
        return np.tile([h ** 3, h ** 3 / 12, h ** 3 / 12, h ** 3 / 12], 3)
    return np.full(3, h ** 3)


def _scatter(op: VieOperator, idxS, Jb):
    nv = int(np.prod(op.dims))
    dense = np.zeros((nv, op.ql), dtype=np.complex128, order="F")
    dense[np.asarray(idxS), :] = np.asarray(Jb).reshape(op.ql, -1).T
    return np.asfortranarray(dense.reshape(op.dims + (op.ql,), order="F"))


def _gather(op: VieOperator, idxS, field):
    nv = int(np.prod(op.dims))
    f = field.reshape(nv, op.ql, order="F")
    return f[np.asarray(idxS), :].T.reshape(-1)


def efield(op: VieOperator, idxS, Jb, e_inc, ce: complex) -> np.ndarray:
    """Total E coefficients at the scatterer voxels (packed)."""
    Jp = _scatter(op, idxS, Jb)
    NJ = op.apply("N", Jp)
    ig = 1.0 / _gram(op)
    eb = (NJ * ig[None, None, None, :] - Jp) / ce
    return np.asarray(e_inc) + _gather(op, idxS, eb)


def bfield(op: VieOperator, idxS, Jb, b_inc) -> np.ndarray:
    """Total B coefficients at the scatterer voxels (packed); needs the K spectrum."""
    Jp = _scatter(op, idxS, Jb)
    KJ = op.apply("K", Jp)
    ig = 1.0 / _gram(op)
    return np.asarray(b_inc) + _gather(op, idxS, KJ * ig[None, None, None, :])
