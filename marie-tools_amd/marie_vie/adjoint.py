"""Adjoint gradient of a field objective w.r.t. voxel permittivity and conductivity.

The reference has no adjoint of the VIE operator and no gradient code (it
only applies the Hermitian adjoint of K inside its basis SVDs,
rSVD_Q_adjoint.m:22-23).  This module composes one from the operator the
library exposes:

  forward   A(theta) J = b(theta),   A = I - D invG N,   b = ce D e_inc,
            D = Mc / Mr = 1 - 1/Mr,  Mr = eps_r + sigma / ce
            (mvp_vie.m:16-17, ie_solver_vie.m:18, em_assembly.m:14-16)
  field     E = J / (ce Mc) in the scatterer (J = ce Mc E)
  objective "e2": f = sum_{v in ROI} sum_c gram_c |E_{c,v}|^2 (integral of
            |E|^2 over the ROI voxels with the Gram weights of mvp_G_*.m);
            "b1": f = sum_{v in ROI} sum_l gram_l |B1+_{l,v}|^2 with
            B = b_inc + invG K J (em_hfield_vie_excitation.m:16-23) and
            B1+ = (Bx + j By) / 2, the co-rotating transmit field;
            "sar": f = sum_{v in ROI} sigma_v / (2 M) sum_c gram_c |E_{c,v}|^2,
            the ROI-average SAR (absorbed power sigma |E|^2 / 2 of the
            e^{+j omega t} phasor, over the ROI mass M = sum rho_v h^3)
  adjoint   A^H lam = dF/dJ*: "e2"/"sar" w J, w = a_v gram_c / |ce Mc|^2 on the
            ROI (a_v = 1, resp. sigma_v / (2 M));
            "b1" L^H P^H W P B, L = gather invG K scatter (L^H uses K^H,
            mvp_K_herm_adj_*_tucker.m), P the B1+ combination, W the ROI Gram
  gradient  df/dtheta_v = 2 Re sum_c conj(lam_cv) u_cv dD_v/dtheta_v
                          (+ for "e2"/"sar" the explicit derivative of the
                          weight a_v / |ce Mc_v|^2 w.r.t. theta_v),
            u = ce e_inc + invG N J = ce e_inc + (J - A J) / D,
            dD/dMr = 1/Mr^2, dMr/deps = 1, dMr/dsigma = 1/ce.
Both solves run on the device GMRES (FP64, or FP32 operator with FP64
refinement); VieObjective wraps it as a torch.autograd.Function.
"""
from __future__ import annotations

import numpy as np

from .operator import VieOperator
from .phantoms import constants


class VieProblem:
    """Fixed geometry, excitation and ROI; the material is the variable."""

    def __init__(self, op: VieOperator, idxS: np.ndarray, freq_hz: float, e_inc: np.ndarray,
                 roi_voxels: np.ndarray, tol: float = 1e-10, precision: str = "fp64",
                 restart: int = 50, maxit: int = 200, kind: str = "e2", b_inc: np.ndarray | None = None,
                 rho: np.ndarray | float = 1000.0, solver: str = "gmres"):
        """kind "e2": |E|^2 over the ROI; kind "b1": |B1+|^2 over the ROI, with
        b_inc the packed incident B coefficients (the operator needs the K
        spectrum); kind "sar": ROI-average SAR with rho the mass density
        (kg/m^3) per scatterer voxel or one value.  solver "gmres" (vie_gmres,
        restart/maxit as is_gmres_vie) or "cocr" (vie_cocr, at most
        restart * maxit iterations) for the forward and the adjoint solve."""
        if solver not in ("gmres", "cocr"):
            raise ValueError("solver must be 'gmres' or 'cocr'")
        self.solver = solver
        if kind not in ("e2", "b1", "sar"):
            raise ValueError("kind must be 'e2', 'b1' or 'sar'")
        if kind == "b1" and b_inc is None:
            raise ValueError("the B1+ objective needs b_inc")
        self.kind = kind
        self.b_inc = None if b_inc is None else np.asarray(b_inc, dtype=np.complex128)
        self.op = op
        self.idxS = np.asarray(idxS, dtype=np.int64)
        self.nS = self.idxS.size
        self.ql = op.ql
        self.ce = constants(freq_hz)["ce"]
        self.e_inc = np.asarray(e_inc, dtype=np.complex128)      # packed coefficients
        roi_voxels = np.asarray(roi_voxels, dtype=np.int64)
        pos = np.minimum(np.searchsorted(self.idxS, roi_voxels), max(self.nS - 1, 0))
        if self.nS == 0 or np.any(self.idxS[pos] != roi_voxels):
            raise ValueError("ROI voxels must lie in the scatterer")
        self.roi = pos                                          # positions in idxS
        h = op.res
        rho = np.broadcast_to(np.asarray(rho, dtype=np.float64), (self.nS,))
        self.roi_mass = float(np.sum(rho[self.roi]) * h ** 3)
        if op.pwl:
            self.gram = np.tile([h ** 3, h ** 3 / 12, h ** 3 / 12, h ** 3 / 12], 3)
        else:
            self.gram = np.full(3, h ** 3)
        self.tol, self.precision, self.restart, self.maxit = tol, precision, restart, maxit

    def _grid(self, vals, fill):
        g = np.full(int(np.prod(self.op.dims)), fill, dtype=np.complex128)
        g[self.idxS] = vals
        return g.reshape(self.op.dims, order="F")

    def set_material(self, eps_r: np.ndarray, sigma: np.ndarray):
        Mr = np.asarray(eps_r, dtype=np.float64) + np.asarray(sigma, dtype=np.float64) / self.ce
        Mc = Mr - 1.0
        self.Mr, self.Mc = Mr, Mc
        self.sigma = np.asarray(sigma, dtype=np.float64)
        self.op.set_material(self._grid(Mr, 1.0), self._grid(Mc, 0.0), self.idxS)
        self.D = Mc / Mr

    def _roi_scale(self):
        """a_v on the ROI: 1 ("e2") or sigma_v / (2 M) ("sar")."""
        if self.kind == "sar":
            return self.sigma[self.roi] / (2.0 * self.roi_mass)
        return np.ones(self.roi.size)

    def weights(self):
        w = np.zeros((self.ql, self.nS))
        inv = self._roi_scale() / np.abs(self.ce * self.Mc[self.roi]) ** 2
        for c in range(self.ql):
            w[c, self.roi] = self.gram[c] * inv
        return w.reshape(-1)

    def _solve(self, b, adjoint=False):
        if self.solver == "cocr":
            return self.op.cocr(b, tol=self.tol, maxit=self.restart * self.maxit, precision=self.precision,
                                adjoint=adjoint)
        return self.op.gmres(b, tol=self.tol, restart=self.restart, maxit=self.maxit, precision=self.precision,
                             adjoint=adjoint)

    def solve(self, eps_r, sigma):
        self.set_material(eps_r, sigma)
        b = self.ce * np.tile(self.D, self.ql) * self.e_inc
        return self._solve(b)

    def _roi_gram(self):
        """W of the B1+ objective: gram_l on the ROI voxels, per (l, voxel)."""
        ls = self.ql // 3
        Wm = np.zeros((ls, self.nS))
        Wm[:, self.roi] = self.gram[:ls, None]
        return Wm

    def b1plus(self, J):
        """B1+ = (Bx + j By) / 2 per (l, voxel), B = b_inc + invG K J."""
        from .fields import bfield
        B = bfield(self.op, self.idxS, J, self.b_inc).reshape(3, self.ql // 3, self.nS)
        return 0.5 * (B[0] + 1j * B[1])

    def objective(self, J) -> float:
        if self.kind == "b1":
            return float(np.sum(self._roi_gram() * np.abs(self.b1plus(J)) ** 2))
        return float(np.sum(self.weights() * np.abs(J) ** 2))

    def adjoint_source(self, J):
        """dF/dJ*: the right-hand side of the adjoint solve."""
        if self.kind in ("e2", "sar"):
            return self.weights() * J
        from .fields import _gather, _gram, _scatter
        t = self._roi_gram() * self.b1plus(J)                # W P B
        q = np.zeros((3, self.ql // 3, self.nS), dtype=np.complex128)
        q[0] = 0.5 * t                                        # P^H: conj of (1/2, j/2, 0)
        q[1] = -0.5j * t
        ig = 1.0 / _gram(self.op)
        Q = _scatter(self.op, self.idxS, q.reshape(-1)) * ig[None, None, None, :]
        return _gather(self.op, self.idxS, self.op.apply("KH", Q))

    def gradient(self, J):
        """(df/deps_r, df/dsigma) on the scatterer voxels, at the current material."""
        lam, info = self._solve(self.adjoint_source(J), adjoint=True)
        AJ = self.op.apply("VIE", J)
        Dp = np.tile(self.D, self.ql)
        u = self.ce * self.e_inc + (J - AJ) / Dp
        s = (np.conj(lam) * u).reshape(self.ql, self.nS).sum(0)
        dD = 1.0 / self.Mr ** 2
        g_eps = 2.0 * np.real(s * dD)
        g_sig = 2.0 * np.real(s * dD / self.ce)
        if self.kind == "b1":        # the weights do not depend on the material
            return g_eps, g_sig, info
        # explicit dependence of the weight a / |ce Mc|^2 on the material:
        # d(1/|Mc|^2)/dtheta = -2 Re(conj(Mc) dMc/dtheta) / |Mc|^4, and for "sar"
        # da/dsigma = 1 / (2 M)
        J2 = (self.gram[:, None] * np.abs(J.reshape(self.ql, self.nS)) ** 2).sum(0)
        a = np.zeros(self.nS)
        a[self.roi] = self._roi_scale()
        base = a * J2 / np.abs(self.ce) ** 2
        Mc = self.Mc
        g_eps += base * (-2.0 * np.real(np.conj(Mc) * 1.0) / np.abs(Mc) ** 4)
        g_sig += base * (-2.0 * np.real(np.conj(Mc) / self.ce) / np.abs(Mc) ** 4)
        if self.kind == "sar":
            mask = np.zeros(self.nS)
            mask[self.roi] = 1.0
            g_sig += mask * J2 / np.abs(self.ce * Mc) ** 2 / (2.0 * self.roi_mass)
        return g_eps, g_sig, info


def _torch_objective():
    import torch

    class VieObjective(torch.autograd.Function):
        """f(eps_r, sigma) with the adjoint-method backward pass."""

        @staticmethod
        def forward(ctx, eps_r, sigma, problem):
            e = eps_r.detach().cpu().numpy()
            s = sigma.detach().cpu().numpy()
            J, _ = problem.solve(e, s)
            ctx.problem = problem
            ctx.J = J
            ctx.dev = eps_r.device
            return torch.tensor(problem.objective(J), dtype=torch.float64, device=eps_r.device)

        @staticmethod
        def backward(ctx, grad_out):
            g_e, g_s, _ = ctx.problem.gradient(ctx.J)
            go = float(grad_out)
            return (torch.as_tensor(go * g_e, device=ctx.dev), torch.as_tensor(go * g_s, device=ctx.dev), None)

    return VieObjective


def vie_objective(eps_r, sigma, problem: VieProblem):
    """torch entry point: differentiable objective of (eps_r, sigma) tensors."""
    return _torch_objective().apply(eps_r, sigma, problem)
