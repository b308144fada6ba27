"""Field recovery and objectives on the device operators.

* fields.efield / fields.bfield (HIP N and K operators) against the oracle's
  restatements of em_efield_vie_excitation.m:18-24 and
  em_hfield_vie_excitation.m:16-23, PWC and PWL, to 1e-12.
* The B1+ objective of the adjoint harness (|(Bx + j By)/2|^2 over a ROI,
  B = b_inc + invG K J) and the ROI-average SAR (sum sigma |E|^2 / (2 M)):
  their adjoint gradients against central finite differences of FP64 device
  solves, B1+ also through torch.autograd; SAR against sigma / (2 M) times
  the |E|^2 objective.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


@pytest.mark.parametrize("pwl", [False, True])
def test_field_recovery_matches_oracle(pwl):
    print(f"case field recovery pwl={pwl}", flush=True)
    from marie_vie import VieOperator, fields
    dims = (9, 8, 7)
    rng = np.random.default_rng(71 + pwl)
    AN = bf.random_kernel(rng, dims, "N", pwl)
    AK = bf.random_kernel(rng, dims, "K", pwl)
    res = 0.005
    ql = 12 if pwl else 3
    nv = int(np.prod(dims))
    idxS = np.sort(rng.choice(nv, size=200, replace=False))
    isol = vo.idx_solution(idxS, nv, ql)
    n = ql * len(idxS)
    Jb = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Ue = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    Ub = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ce = vo.em_constants(298e6)["ce"]
    op = VieOperator(dims, "PWL" if pwl else "PWC", res)
    try:
        op.set_green_dense("N", AN)
        op.set_green_dense("K", AK)
        E = fields.efield(op, idxS, Jb, Ue, ce)
        B = fields.bfield(op, idxS, Jb, Ub)
    finally:
        op.close()
    TN = vo.circulant_dense(AN, pwl)
    TK = vo.circulant_dense(AK, pwl)
    Eref = vo.efield_vie(Jb, Ue, ce, isol, dims, res, TN, pwl)
    Bref = vo.hfield_vie(Jb, Ub, isol, dims, res, TK, pwl)
    e1, e2 = _rel(E, Eref), _rel(B, Bref)
    print(f"field recovery pwl={pwl}: E {e1:.2e}, B {e2:.2e}", flush=True)
    assert e1 < 1e-12 and e2 < 1e-12


def _b1_problem():
    from marie_vie import VieOperator, green, phantoms
    from marie_vie.adjoint import VieProblem
    freq = 298e6
    ph = phantoms.sphere(dims=(10, 10, 10), res=0.01, radius=0.045, eps_r=50.0, sigma=0.6)
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWC", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    op.set_green_dense("K", green.assemble(op, "K", const["k0"]))
    e_inc = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
    # incident B of the x-polarised plane wave along +z: (k0 / omega) y_hat exp(-j k0 z)
    bscale = const["k0"] / const["omega"]
    b_inc = phantoms.project(ph, lambda x, y, z: (0 * z, bscale * np.exp(-1j * const["k0"] * z), 0 * z), False)
    # ROI: the voxels of the central 4x4x4 cube
    c = np.arange(3, 7)
    I, J, K = np.meshgrid(c, c, c, indexing="ij")
    roi = np.sort((I + 10 * (J + 10 * K)).reshape(-1))
    prob = VieProblem(op, ph.idxS, freq, e_inc, roi, tol=1e-12, precision="fp64", kind="b1", b_inc=b_inc)
    eps = ph.eps_r.reshape(-1, order="F")[ph.idxS]
    sig = ph.sigma.reshape(-1, order="F")[ph.idxS]
    return op, prob, eps, sig


def test_b1_objective_gradient_matches_finite_differences():
    print("case B1+ gradient vs FD", flush=True)
    op, prob, eps, sig = _b1_problem()
    try:
        J, info = prob.solve(eps, sig)
        f0 = prob.objective(J)
        assert f0 > 0
        g_eps, g_sig, _ = prob.gradient(J)
        rng = np.random.default_rng(3001)
        pick = rng.choice(prob.roi, size=4, replace=False)
        h = 1e-4
        worst = 0.0
        for which in ("eps", "sig"):
            for v in pick:
                fs = []
                for sgn in (1.0, -1.0):
                    ep, sp = eps.copy(), sig.copy()
                    (ep if which == "eps" else sp)[v] += sgn * h
                    Jp, _ = prob.solve(ep, sp)
                    fs.append(prob.objective(Jp))
                fd = (fs[0] - fs[1]) / (2 * h)
                g = g_eps[v] if which == "eps" else g_sig[v]
                err = abs(g - fd) / abs(fd)
                worst = max(worst, err)
                print(f"B1+ {which} voxel {v}: adjoint {g:.6e} fd {fd:.6e} rel {err:.1e}", flush=True)
        assert worst < 1e-4
    finally:
        op.close()


def test_b1_objective_through_torch_autograd():
    print("case B1+ torch autograd", flush=True)
    import torch
    from marie_vie.adjoint import vie_objective
    op, prob, eps, sig = _b1_problem()
    try:
        te = torch.tensor(eps, dtype=torch.float64, device="cuda", requires_grad=True)
        ts = torch.tensor(sig, dtype=torch.float64, device="cuda", requires_grad=True)
        f = vie_objective(te, ts, prob)
        f.backward()
        J, _ = prob.solve(eps, sig)
        g_eps, g_sig, _ = prob.gradient(J)
        assert np.allclose(te.grad.cpu().numpy(), g_eps, rtol=1e-9, atol=1e-9 * np.abs(g_eps).max())
        assert np.allclose(ts.grad.cpu().numpy(), g_sig, rtol=1e-9, atol=1e-9 * np.abs(g_sig).max())
    finally:
        op.close()


def _sar_problem(kind="sar"):
    from marie_vie import VieOperator, green, phantoms
    from marie_vie.adjoint import VieProblem
    freq = 298e6
    ph = phantoms.sphere(dims=(10, 10, 10), res=0.01, radius=0.045, eps_r=50.0, sigma=0.6)
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWC", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    e_inc = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
    c = np.arange(3, 7)
    I, J, K = np.meshgrid(c, c, c, indexing="ij")
    roi = np.sort((I + 10 * (J + 10 * K)).reshape(-1))
    rho = 1000.0 + 100.0 * np.random.default_rng(5).random(len(ph.idxS))   # kg/m^3 per voxel
    prob = VieProblem(op, ph.idxS, freq, e_inc, roi, tol=1e-12, precision="fp64", kind=kind, rho=rho)
    eps = ph.eps_r.reshape(-1, order="F")[ph.idxS]
    sig = ph.sigma.reshape(-1, order="F")[ph.idxS]
    return op, prob, eps, sig, rho


def test_sar_objective_is_sigma_weighted_field_energy():
    """With a uniform sigma on the ROI, the ROI-average SAR equals
    sigma / (2 M) times the |E|^2 objective (M the ROI mass)."""
    print("case SAR = sigma / 2M |E|^2", flush=True)
    from marie_vie.adjoint import VieProblem
    op, prob, eps, sig, rho = _sar_problem()
    try:
        J, _ = prob.solve(eps, sig)
        e2 = VieProblem(op, prob.idxS, 298e6, prob.e_inc, prob.idxS[prob.roi], kind="e2")
        e2.set_material(eps, sig)
        M = float(np.sum(rho[prob.roi]) * op.res ** 3)
        want = 0.6 / (2 * M) * e2.objective(J)
        assert abs(prob.objective(J) - want) <= 1e-13 * want
    finally:
        op.close()


def test_sar_objective_gradient_matches_finite_differences():
    print("case SAR gradient vs FD", flush=True)
    op, prob, eps, sig, _ = _sar_problem()
    try:
        J, _ = prob.solve(eps, sig)
        g_eps, g_sig, _ = prob.gradient(J)
        rng = np.random.default_rng(3002)
        pick = rng.choice(prob.roi, size=4, replace=False)
        h = 1e-4
        worst = 0.0
        for which in ("eps", "sig"):
            for v in pick:
                fs = []
                for sgn in (1.0, -1.0):
                    ep, sp = eps.copy(), sig.copy()
                    (ep if which == "eps" else sp)[v] += sgn * h
                    Jp, _ = prob.solve(ep, sp)
                    fs.append(prob.objective(Jp))
                fd = (fs[0] - fs[1]) / (2 * h)
                g = g_eps[v] if which == "eps" else g_sig[v]
                err = abs(g - fd) / abs(fd)
                worst = max(worst, err)
                print(f"SAR {which} voxel {v}: adjoint {g:.6e} fd {fd:.6e} rel {err:.1e}", flush=True)
        assert worst < 1e-4
    finally:
        op.close()
