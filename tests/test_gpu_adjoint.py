"""Adjoint gradient (device GMRES on A and A^H) against central finite
differences computed entirely by the CPU oracle (dense direct solves of the
oracle's VIE matrix), on a small PWC and PWL problem with the assembled
physical Green tensor.  FD step 1e-4 (absolute), as the C3 workload states."""
import numpy as np
import pytest

from oracle import vie_oracle as vo

pytestmark = pytest.mark.gpu


def _problem(pwl, precision="fp64", tol=1e-12):
    from marie_vie import VieOperator, green, phantoms
    from marie_vie.adjoint import VieProblem
    dims = (7, 6, 8)
    res = 0.01
    freq = 298e6
    ph = phantoms.ellipsoid_regions(dims=dims, res=res, semi=(0.03, 0.025, 0.035), ncell=3,
                                    eps_range=(40, 80), sig_range=(0.3, 0.9), seed=3003)
    const = phantoms.constants(freq)
    op = VieOperator(dims, "PWL" if pwl else "PWC", res)
    A = green.assemble(op, "N", const["k0"])
    op.set_green_dense("N", A)
    e_inc = phantoms.project(ph, phantoms.plane_wave(const["k0"]), pwl)
    rng = np.random.default_rng(3004)
    roi = np.sort(rng.choice(ph.idxS, size=6, replace=False))
    prob = VieProblem(op, ph.idxS, freq, e_inc, roi, tol=tol, precision=precision)
    eps = ph.eps_r.reshape(-1, order="F")[ph.idxS]
    sig = ph.sigma.reshape(-1, order="F")[ph.idxS]
    return prob, ph, A, eps, sig, const


_NMAT = {}


def _oracle_f(prob, ph, A, eps, sig, const, pwl):
    """f computed by the oracle alone: dense VIE matrix, direct solve.
    The VIE matrix is I - diag(Mc/Mr) invG Ns with Ns the oracle's N operator
    restricted to the scatterer (built once from unit vectors; material-free)."""
    dims = ph.dims
    nv = int(np.prod(dims))
    ql = 12 if pwl else 3
    isol = vo.idx_solution(ph.idxS, nv, ql)
    n = len(isol)
    key = (pwl, dims)
    if key not in _NMAT:
        T = vo.circulant_dense(A if pwl else A[:, :, :, 0, :], pwl)
        Ns = np.zeros((n, n), dtype=complex)
        for k in range(n):
            Jp = np.zeros(nv * ql, dtype=complex)
            Jp[isol[k]] = 1.0
            NJ = vo.mvp_N(Jp.reshape(dims + (ql,), order="F"), T, pwl)
            Ns[:, k] = NJ.reshape(-1, order="F")[isol]
        _NMAT[key] = Ns
    Ns = _NMAT[key]
    Mr_s = eps + sig / const["ce"]
    Mr = np.ones(nv, dtype=complex)
    Mr[ph.idxS] = Mr_s
    Mr = Mr.reshape(dims, order="F")
    Mc = Mr - 1.0
    Dp = np.tile(Mc.reshape(-1, order="F")[ph.idxS] / Mr_s, ql)
    ig = np.repeat(1.0 / vo.gram_diag(pwl, ph.res), len(ph.idxS))
    M = np.eye(n) - (Dp * ig)[:, None] * Ns
    b = vo.rhs_vie(const["ce"], Mr, Mc, ph.idxS, ql, prob.e_inc)
    J = np.linalg.solve(M, b).reshape(ql, -1)
    Mcs = Mc.reshape(-1, order="F")[ph.idxS]
    E2 = (prob.gram[:, None] * np.abs(J / (const["ce"] * Mcs)) ** 2)
    return float(E2[:, prob.roi].sum())


@pytest.mark.parametrize("pwl", [False, True])
def test_adjoint_gradient_vs_oracle_finite_differences(pwl):
    print(f"case adjoint pwl={pwl}", flush=True)
    prob, ph, A, eps, sig, const = _problem(pwl)
    J, info = prob.solve(eps, sig)
    f_gpu = prob.objective(J)
    f_or = _oracle_f(prob, ph, A, eps, sig, const, pwl)
    assert abs(f_gpu - f_or) / abs(f_or) < 1e-9
    g_eps, g_sig, _ = prob.gradient(J)
    rng = np.random.default_rng(3005)
    pick = rng.choice(len(eps), size=6, replace=False)
    h = 1e-4
    for v in pick:
        for which in ("eps", "sig"):
            ep, sp = eps.copy(), sig.copy()
            em, sm = eps.copy(), sig.copy()
            if which == "eps":
                ep[v] += h
                em[v] -= h
            else:
                sp[v] += h
                sm[v] -= h
            fd = (_oracle_f(prob, ph, A, ep, sp, const, pwl) - _oracle_f(prob, ph, A, em, sm, const, pwl)) / (2 * h)
            g = g_eps[v] if which == "eps" else g_sig[v]
            scale = max(abs(fd), 1e-3 * np.abs(g_sig if which == "sig" else g_eps).max())
            assert abs(g - fd) / scale < 1e-4, (which, v, g, fd)
    prob.op.close()


def test_mixed_precision_gradient_close_to_fp64():
    prob64, ph, A, eps, sig, const = _problem(True, "fp64", 1e-10)
    J64, _ = prob64.solve(eps, sig)
    ge64, gs64, _ = prob64.gradient(J64)
    prob64.precision = "mixed"
    prob64.tol = 1e-8
    J32, _ = prob64.solve(eps, sig)
    ge32, gs32, _ = prob64.gradient(J32)
    assert np.linalg.norm(J32 - J64) / np.linalg.norm(J64) < 1e-6
    assert np.linalg.norm(ge32 - ge64) / np.linalg.norm(ge64) < 1e-5
    assert np.linalg.norm(gs32 - gs64) / np.linalg.norm(gs64) < 1e-5
    prob64.op.close()


def test_cocr_solver_gradient_matches_gmres():
    """The harness with the COCR solver for the forward and the adjoint
    (VIE^H) solve: FP64 gradient equal to the GMRES one, and the mixed COCR
    gradient within 1e-5 of it."""
    prob, ph, A, eps, sig, const = _problem(True, "fp64", 1e-10)
    J, _ = prob.solve(eps, sig)
    ge, gs, _ = prob.gradient(J)
    prob.solver = "cocr"
    Jc, _ = prob.solve(eps, sig)
    gec, gsc, info = prob.gradient(Jc)
    assert info["nreplace"][0] >= 1
    assert np.linalg.norm(Jc - J) / np.linalg.norm(J) < 1e-8
    assert np.linalg.norm(gec - ge) / np.linalg.norm(ge) < 1e-7
    assert np.linalg.norm(gsc - gs) / np.linalg.norm(gs) < 1e-7
    prob.precision = "mixed"
    prob.tol = 1e-8
    Jm, _ = prob.solve(eps, sig)
    gem, gsm, _ = prob.gradient(Jm)
    assert np.linalg.norm(gem - ge) / np.linalg.norm(ge) < 1e-5
    assert np.linalg.norm(gsm - gs) / np.linalg.norm(gs) < 1e-5
    prob.op.close()


def test_torch_autograd_wrapper():
    import torch
    from marie_vie.adjoint import vie_objective
    prob, ph, A, eps, sig, const = _problem(False, "fp64", 1e-11)
    e = torch.tensor(eps, dtype=torch.float64, requires_grad=True)
    s = torch.tensor(sig, dtype=torch.float64, requires_grad=True)
    f = vie_objective(e, s, prob)
    f.backward()
    J, _ = prob.solve(eps, sig)
    ge, gs, _ = prob.gradient(J)
    np.testing.assert_allclose(e.grad.numpy(), ge, rtol=1e-8, atol=1e-14 * np.abs(ge).max())
    np.testing.assert_allclose(s.grad.numpy(), gs, rtol=1e-8, atol=1e-14 * np.abs(gs).max())
    prob.op.close()


def test_adjoint_gradient_at_c3_scale():
    """The C3 workload (64x64x80 PWL head at 298 MHz, 8x8x8 ROI): adjoint
    gradient against central differences of FP64 device solves (tol 1e-12,
    step 1e-4) on 4 seeded ROI voxels, and the mixed-precision gradient
    against the FP64 one (bench.py --adjoint runs 32 points)."""
    import bench
    r = bench.adjoint_section(nfd=4)
    print(f"C3 adjoint: fd rel err eps {r['fd_max_relerr_eps']:.2e} sigma {r['fd_max_relerr_sigma']:.2e}, "
          f"mixed vs fp64 {r['mixed_vs_fp64_grad_relerr_eps']:.2e}", flush=True)
    assert r["fd_max_relerr_eps"] < 1e-4 and r["fd_max_relerr_sigma"] < 1e-4
    assert r["mixed_vs_fp64_grad_relerr_eps"] < 1e-5 and r["mixed_vs_fp64_grad_relerr_sigma"] < 1e-5
