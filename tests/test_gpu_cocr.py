"""GPU tests of the device COCR solver (vie_cocr, through the C ABI).

* W-symmetry of the device VIE / VIE^H operators (FP64) on random-kernel
  systems and on the C1 sphere with its assembled Green tensor: the premise
  of the recurrence holds for the operator the library applies.
* FP64 COCR equals the numpy recurrence of tests/krylov_ref.py (itself
  pinned against dense solves in test_cocr_ref.py) iterate by iterate:
  resvec to 1e-6 relative, the same number of operator applications, x to
  1e-9; and x equals the oracle's dense solve.
* Mixed form: the FP64 residual of the returned x (oracle operator) is below
  tol, and x is within 10 tol of the FP64 solution.
* Batched equals single solves; graphs equal eager launches bitwise; warm
  start; maxit; the adjoint system.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf
import krylov_ref as kr

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _system(seed, dims, pwl, nrhs=1, frac=0.6, res=0.005, scale=0.6):
    rng = np.random.default_rng(seed)
    A = bf.random_kernel(rng, dims, "N", pwl)
    nv = int(np.prod(dims))
    eps = 1 + 60 * rng.random(dims)
    sig = rng.random(dims)
    idxS = np.sort(rng.choice(nv, size=max(1, int(frac * nv)), replace=False))
    mat = vo.material(eps, sig, vo.em_constants(298e6)["ce"])
    ql = 12 if pwl else 3
    isol = vo.idx_solution(idxS, nv, ql)
    n = len(isol)
    B = rng.standard_normal((n, nrhs)) + 1j * rng.standard_normal((n, nrhs))
    T0 = vo.circulant_dense(A, pwl)
    d = B[:, 0] - vo.mvp_vie(B[:, 0], mat["Mr"], mat["Mc"], isol, dims, res, T0, pwl)
    A = A * (scale * np.linalg.norm(B[:, 0]) / np.linalg.norm(d))      # compact part of norm ~scale
    T = vo.circulant_dense(A, pwl)
    fwd = lambda v: vo.mvp_vie(v, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    adj = lambda v: vo.mvp_vie_adjoint(v, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    return dict(A=A, mat=mat, idxS=idxS, B=np.asfortranarray(B), fwd=fwd, adj=adj, ql=ql, res=res, pwl=pwl,
                dims=dims)


def _op(s):
    from marie_vie import VieOperator
    op = VieOperator(s["dims"], "PWL" if s["pwl"] else "PWC", s["res"])
    op.set_green_dense("N", s["A"])
    op.set_material(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"])
    return op


def _w(s, adjoint=False):
    return kr.vie_weights(s["mat"]["Mr"], s["mat"]["Mc"], s["idxS"], s["ql"], s["res"], s["pwl"], adjoint)


@pytest.mark.parametrize("pwl", [False, True])
def test_device_operator_symmetric_in_weighted_form(pwl):
    print(f"case pwl={pwl}", flush=True)
    s = _system(41 + pwl, (7, 6, 5), pwl)
    op = _op(s)
    rng = np.random.default_rng(9)
    n = s["B"].shape[0]
    U = np.asfortranarray(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))
    for name, adjoint in (("VIE", False), ("VIEH", True)):
        w = _w(s, adjoint)
        AU = op.apply(name, U)
        lhs, rhs = kr.bil(AU[:, 0], U[:, 1], w), kr.bil(U[:, 0], AU[:, 1], w)
        assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    op.close()


def test_sphere_green_tensor_operator_symmetric():
    """The assembled Green tensor (far/medium quadrature + near field) of the
    C1-type sphere keeps the symmetry."""
    print("case sphere 16^3", flush=True)
    from marie_vie import VieOperator, green, phantoms
    ph = phantoms.sphere(dims=(16, 16, 16), res=0.01)
    const = phantoms.constants(298e6)
    op = VieOperator(ph.dims, "PWC", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(298e6)
    op.set_material(Mr, Mc, ph.idxS)
    w = kr.vie_weights(Mr, Mc, ph.idxS, 3, ph.res, False)
    rng = np.random.default_rng(10)
    n = 3 * len(ph.idxS)
    U = np.asfortranarray(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))
    AU = op.apply("VIE", U)
    lhs, rhs = kr.bil(AU[:, 0], U[:, 1], w), kr.bil(U[:, 0], AU[:, 1], w)
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
def test_fp64_cocr_matches_numpy_recurrence(pwl):
    print(f"case pwl={pwl}", flush=True)
    s = _system(51 + pwl, (6, 5, 7), pwl)
    op = _op(s)
    b = s["B"][:, 0]
    x, info = op.cocr(b, tol=1e-10, maxit=400)
    xr, rvr, nmvr = kr.cocr(s["fwd"], b, _w(s), 1e-10, 400)
    assert info["nmatvec"][0] == nmvr and info["nreplace"][0] == 1
    assert len(info["resvec"][0]) == len(rvr)
    np.testing.assert_allclose(info["resvec"][0], rvr, rtol=1e-6, atol=1e-14)
    assert _rel(x, xr) < 1e-9
    assert np.linalg.norm(b - s["fwd"](x)) < 1e-10 * np.linalg.norm(b)
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
def test_mixed_cocr_meets_fp64_residual(pwl):
    print(f"case pwl={pwl}", flush=True)
    s = _system(61 + pwl, (6, 5, 7), pwl)
    op = _op(s)
    b = s["B"][:, 0]
    tol = 1e-8
    x, info = op.cocr(b, tol=tol, maxit=400, precision="mixed")
    assert np.linalg.norm(b - s["fwd"](x)) < tol * np.linalg.norm(b)     # the oracle's FP64 operator
    assert info["resvec"][0][-1] < tol and info["nreplace"][0] >= 2      # replacements on the way
    x64, _ = op.cocr(b, tol=1e-12, maxit=400)
    assert _rel(x, x64) < 10 * tol
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_batched_cocr_equals_single(precision):
    """Three rhs in lockstep, retired at different iterations: each equals its
    own single solve (bitwise for FP64; the mixed form within tol)."""
    s = _system(71, (6, 5, 7), False, nrhs=3)
    s["B"][:, 1] *= 1e-3                     # another scale: same relative path
    s["B"][:, 2] = s["B"][:, 0] + 0.5 * s["B"][:, 2]
    op = _op(s)
    tol = 1e-9
    X, info = op.cocr(s["B"], tol=tol, maxit=400, precision=precision)
    for k in range(3):
        x1, i1 = op.cocr(np.ascontiguousarray(s["B"][:, k]), tol=tol, maxit=400, precision=precision)
        if precision == "fp64":
            np.testing.assert_array_equal(X[:, k], x1)
            assert info["nmatvec"][k] == i1["nmatvec"][0]
        else:
            assert _rel(X[:, k], x1) < 10 * tol
        assert np.linalg.norm(s["B"][:, k] - s["fwd"](X[:, k])) < tol * np.linalg.norm(s["B"][:, k])
    op.close()


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cocr_graphs_match_eager(precision, monkeypatch):
    """Blocks of iterations run as captured graphs (replayed across blocks
    and solves); VIE_GRAPHS=0 launches the same kernels eagerly.  Every
    replay, with eager work in between, equals the eager solve bitwise."""
    s = _system(81, (6, 5, 7), False, nrhs=2)
    monkeypatch.setenv("VIE_AXIS_FFT", "ct")
    res = {}
    for graphs in ("0", "1"):
        monkeypatch.setenv("VIE_GRAPHS", graphs)
        op = _op(s)
        runs = []
        for rep in range(3):
            runs.append(op.cocr(s["B"], tol=1e-10, maxit=400, precision=precision))
            op.apply("VIE", np.ascontiguousarray(s["B"][:, 0]))
        op.close()
        res[graphs] = runs
    x0, i0 = res["0"][0]
    assert min(i0["nmatvec"]) > 20                      # more than one block of 16
    for graphs in ("0", "1"):
        for x, info in res[graphs]:
            np.testing.assert_array_equal(x, x0)
            assert info["nmatvec"] == i0["nmatvec"] and info["nreplace"] == i0["nreplace"]
            for a, b in zip(info["resvec"], i0["resvec"]):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_cocr_stall_restarts_still_converge(precision, monkeypatch):
    """A stall window of 2 iterations forces the restart path (replacement,
    recurrence restarted from the FP64 residual) many times: the solve still
    ends on the FP64 residual below tol, with more replacements than without."""
    s = _system(95, (6, 5, 7), True, scale=0.9)        # slow enough to stall a 2-iteration window
    b = s["B"][:, 0]
    op = _op(s)
    x0, i0 = op.cocr(b, tol=1e-8, maxit=2000, precision=precision)
    op.close()
    monkeypatch.setenv("VIE_COCR_STALL", "2")
    op = _op(s)
    x, info = op.cocr(b, tol=1e-8, maxit=2000, precision=precision)
    op.close()
    assert info["nreplace"][0] > i0["nreplace"][0]
    assert np.linalg.norm(b - s["fwd"](x)) < 1e-8 * np.linalg.norm(b)
    assert _rel(x, x0) < 1e-6


def test_cocr_replacement_factor(monkeypatch):
    """The mixed form replaces its recurrence residual by the FP64 residual
    whenever it has fallen by VIE_COCR_DELTA (default 0.1; tested between
    graphs of 16 iterations, so the system converges slowly): a larger factor
    replaces more often, a smaller one no more often, and every setting ends on the
    FP64 residual below tol."""
    s = _system(95, (6, 5, 7), True, scale=0.9)        # slow: many iterations per decade
    b = s["B"][:, 0]
    nrep = {}
    for delta in ("0.5", "", "0.001"):
        if delta:
            monkeypatch.setenv("VIE_COCR_DELTA", delta)
        else:
            monkeypatch.delenv("VIE_COCR_DELTA", raising=False)
        op = _op(s)
        x, info = op.cocr(b, tol=1e-8, maxit=2000, precision="mixed")
        op.close()
        assert np.linalg.norm(b - s["fwd"](x)) < 1e-8 * np.linalg.norm(b)
        nrep[delta] = info["nreplace"][0]
    print(f"replacements by factor {nrep}", flush=True)
    assert nrep["0.5"] > nrep[""] >= nrep["0.001"], nrep


def test_cocr_warm_start_maxit_and_adjoint():
    s = _system(91, (6, 5, 7), True)
    op = _op(s)
    b = s["B"][:, 0]
    xr, _, _ = kr.cocr(s["fwd"], b, _w(s), 1e-12, 600)
    # warm start near the solution: fewer iterations, same solution
    x0 = xr + 1e-4 * np.abs(xr).max() * np.random.default_rng(3).standard_normal(len(b))
    x, info = op.cocr(b, x0=x0, tol=1e-10, maxit=400)
    xc, infc = op.cocr(b, tol=1e-10, maxit=400)
    assert info["nmatvec"][0] < infc["nmatvec"][0]
    assert _rel(x, xr) < 1e-8
    # maxit: stops after maxit iterations with the FP64 residual of x last
    xm, im = op.cocr(b, tol=1e-14, maxit=5)
    assert im["nmatvec"][0] == 6 and len(im["resvec"][0]) == 7
    assert abs(im["resvec"][0][-1] - np.linalg.norm(b - s["fwd"](xm)) / np.linalg.norm(b)) < 1e-12
    # the adjoint system (weights conj(d))
    xa, ia = op.cocr(b, tol=1e-10, maxit=400, adjoint=True)
    xar, rva, nmva = kr.cocr(s["adj"], b, _w(s, adjoint=True), 1e-10, 400)
    assert ia["nmatvec"][0] == nmva and _rel(xa, xar) < 1e-9
    op.close()


def test_c1_sphere_cocr_agrees_with_gmres():
    """The C1 workload (32^3 PWC sphere, 298 MHz): FP64 COCR and FP64 GMRES
    solutions to 1e-10 agree; the mixed COCR solve to 1e-6 meets the FP64
    residual and lies within 1e-5 of them."""
    print("case c1 sphere", flush=True)
    import torch
    from marie_vie import VieOperator, green, phantoms
    ph = phantoms.sphere()
    const = phantoms.constants(298e6)
    op = VieOperator(ph.dims, "PWC", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(298e6)
    op.set_material(Mr, Mc, ph.idxS)
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    b = const["ce"] * np.tile(mcr, 3) * phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
    bt = torch.from_numpy(b.reshape(1, -1).copy()).cuda()
    xg, ig = op.gmres(bt, tol=1e-10, restart=50, maxit=200)
    xc, ic = op.cocr(bt, tol=1e-10, maxit=2000)
    xm, im = op.cocr(bt, tol=1e-6, maxit=2000, precision="mixed")
    r = lambda x: float(torch.linalg.vector_norm(bt - op.apply("VIE", x.contiguous())) / torch.linalg.vector_norm(bt))
    rel = lambda a, b: float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(b))
    assert r(xc) < 1e-10 and r(xm) < 1e-6
    assert rel(xc, xg) < 1e-8
    assert rel(xm, xg) < 1e-5
    print(f"c1: gmres {ig['nmatvec'][0]} cocr {ic['nmatvec'][0]} mixed cocr {im['nmatvec'][0]} "
          f"(+{im['nreplace'][0]} fp64) matvecs", flush=True)
    op.close()
