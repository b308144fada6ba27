"""Edge cases of the operators and the solver on the device, against the oracle.

* Grids with a single voxel along one or more axes (FFT length 2), which the
  reference's assembly and embedding handle (assembly_N.m:59-66; the
  embedding [U; 0] of length 2): N and K, PWC and PWL, forward and Hermitian
  adjoint, dense spectrum and the same Tucker T-struct on both sides; the
  packed VIE operator on them.  The oracle itself is checked on these grids
  against brute-force Toeplitz matrices (tests/test_oracle.py).
* A one-voxel scatterer with the support box on: the packed operators then
  run on a 1 x 1 x 1 child context (FFT 2 x 2 x 2); VIE and VIE^H in FP64 and
  FP32, and GMRES ([p outit], resvec, x) against the oracle's is_gmres_vie.
* An empty scatterer (nS = 0): the packed operators do nothing and succeed,
  GMRES refuses with VIE_ERR_ARG.
* Batch limit: nrhs beyond 65535 / ql is refused before any launch.
"""
import ctypes

import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf

pytestmark = pytest.mark.gpu

TOL64, TOL32 = 1e-12, 2e-5
DEGENERATE = [(1, 5, 4), (4, 1, 3), (3, 4, 1), (1, 1, 5), (6, 1, 1), (1, 1, 1)]


def _rel(a, b):
    """Max-abs relative difference; against an exactly zero reference (K on a
    one-voxel PWC grid: its antisymmetric kernel has no self term) the size
    of a itself."""
    s = np.abs(b).max()
    return np.abs(a - b).max() / s if s > 0 else np.abs(a).max()


def _op(dims, pwl, res=0.005):
    from marie_vie import VieOperator
    return VieOperator(dims, "PWL" if pwl else "PWC", res)


@pytest.mark.parametrize("dims", DEGENERATE)
@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("opname", ["N", "K"])
def test_single_voxel_axes_match_oracle(dims, pwl, opname):
    print(f"case degenerate grid dims={dims} pwl={pwl} op={opname}", flush=True)
    rng = np.random.default_rng(80 + sum(dims) + 7 * pwl + (opname == "K"))
    A = bf.random_kernel(rng, dims, opname, pwl)
    ql = 12 if pwl else 3
    J = np.asfortranarray(rng.standard_normal(dims + (ql,)) + 1j * rng.standard_normal(dims + (ql,)))
    apply = vo.mvp_N if opname == "N" else vo.mvp_K
    for form in ("dense", "tucker"):
        op = _op(dims, pwl)
        try:
            if form == "dense":
                op.set_green_dense(opname, A)
            T = vo.circulant_dense(A, pwl) if form == "dense" else \
                (vo.fft_circ_tucker_pwl if pwl else vo.fft_circ_tucker_pwc)(A, None)
            if form == "tucker":
                op.set_green_tucker(opname, T)
            got = op.apply(opname, J)
            gotH = op.apply(opname + "H", J)
        finally:
            op.close()
        assert _rel(got, apply(J, T, pwl)) < TOL64, form
        assert _rel(gotH, apply(J, T, pwl, adjoint=True)) < TOL64, form


@pytest.mark.parametrize("dims", [(1, 5, 4), (3, 4, 1), (1, 1, 5)])
@pytest.mark.parametrize("pwl", [False, True])
def test_vie_operator_on_single_voxel_axes(dims, pwl):
    print(f"case degenerate grid VIE dims={dims} pwl={pwl}", flush=True)
    rng = np.random.default_rng(90 + sum(dims) + pwl)
    A = bf.random_kernel(rng, dims, "N", pwl)
    nv = int(np.prod(dims))
    idxS = np.sort(rng.choice(nv, size=max(1, (2 * nv) // 3), replace=False))
    mat = vo.material(1 + 40 * rng.random(dims), rng.random(dims), vo.em_constants(298e6)["ce"])
    ql = 12 if pwl else 3
    x = rng.standard_normal(ql * len(idxS)) + 1j * rng.standard_normal(ql * len(idxS))
    op = _op(dims, pwl)
    try:
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        y = op.apply("VIE", x)
        yh = op.apply("VIEH", x)
        y32 = op.apply("VIE", x, precision="fp32")
    finally:
        op.close()
    isol = vo.idx_solution(idxS, nv, ql)
    T = vo.circulant_dense(A, pwl)
    ref = vo.mvp_vie(x, mat["Mr"], mat["Mc"], isol, dims, 0.005, T, pwl)
    refh = vo.mvp_vie_adjoint(x, mat["Mr"], mat["Mc"], isol, dims, 0.005, T, pwl)
    assert _rel(y, ref) < TOL64 and _rel(yh, refh) < TOL64 and _rel(y32, ref) < TOL32


@pytest.mark.parametrize("pwl", [False, True])
def test_one_voxel_scatterer_on_its_support_box(pwl):
    """One scatterer voxel in a 10 x 9 x 8 grid: the support box is 1 x 1 x 1
    (a child context with FFT 2 x 2 x 2); the operators and GMRES match the
    oracle on the full grid."""
    print(f"case one-voxel scatterer pwl={pwl}", flush=True)
    dims = (10, 9, 8)
    rng = np.random.default_rng(95 + pwl)
    A = bf.random_kernel(rng, dims, "N", pwl)
    nv = int(np.prod(dims))
    v = 4 + dims[0] * (3 + dims[1] * 5)
    idxS = np.array([v])
    mat = vo.material(1 + 60 * rng.random(dims), rng.random(dims), vo.em_constants(298e6)["ce"])
    ql = 12 if pwl else 3
    x = rng.standard_normal(ql) + 1j * rng.standard_normal(ql)
    b = rng.standard_normal(ql) + 1j * rng.standard_normal(ql)
    op = _op(dims, pwl)
    try:
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        box = op.box()
        assert box is not None and box["extent"] == (1, 1, 1) and box["fft"] == (2, 2, 2), box
        y = op.apply("VIE", x)
        yh = op.apply("VIEH", x)
        y32 = op.apply("VIEH", x, precision="fp32")
        xs, info = op.gmres(b, tol=1e-6, restart=50, maxit=200)
    finally:
        op.close()
    isol = vo.idx_solution(idxS, nv, ql)
    T = vo.circulant_dense(A, pwl)
    afun = lambda u: vo.mvp_vie(u, mat["Mr"], mat["Mc"], isol, dims, 0.005, T, pwl)
    assert _rel(y, afun(x)) < TOL64
    refh = vo.mvp_vie_adjoint(x, mat["Mr"], mat["Mc"], isol, dims, 0.005, T, pwl)
    assert _rel(yh, refh) < TOL64 and _rel(y32, refh) < TOL32
    xr, itr, rvr = vo.gmres_vie(afun, b, restart=50, tol=1e-6, maxit=200)
    print(f"one voxel GMRES: device {info['iters'][0]}, oracle {list(itr)}", flush=True)
    assert info["iters"][0] == list(itr)
    np.testing.assert_allclose(info["resvec"][0], rvr, rtol=1e-6, atol=1e-14)
    assert _rel(xs, xr) < 1e-8


def test_empty_scatterer_and_batch_limit():
    print("case empty scatterer, batch limit", flush=True)
    from marie_vie import _lib as L
    dims = (6, 5, 4)
    rng = np.random.default_rng(97)
    A = bf.random_kernel(rng, dims, "N", False)
    mat = vo.material(1 + 10 * rng.random(dims), rng.random(dims), vo.em_constants(298e6)["ce"])
    op = _op(dims, False)
    try:
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], np.zeros(0, dtype=np.int64))
        buf = np.zeros(4, dtype=np.complex128)
        out = np.full(4, 7.0 + 0j)
        p = lambda a: ctypes.c_void_p(a.ctypes.data)
        for code in (L.VIE_OP_VIE, L.VIE_OP_VIEH):
            assert op.lib.vie_apply(op.h, code, L.VIE_PREC_FP64, p(buf), p(out), 1, L.VIE_MEM_HOST) == L.VIE_OK
        assert np.all(out == 7.0)                      # nothing written
        rv = np.zeros(8)
        nres = np.zeros(1, dtype=np.int32)
        its = np.zeros(2, dtype=np.int32)
        nmv = np.zeros(1, dtype=np.int32)
        rc = op.lib.vie_gmres(op.h, L.VIE_OP_VIE, L.VIE_PREC_FP64, p(buf), p(out), 1, 1e-6, 5, 1,
                              L.VIE_MEM_HOST, p(rv), p(nres), p(its), p(nmv))
        assert rc == L.VIE_ERR_ARG
        # nrhs beyond the launch grid's limit (grid.y = rhs x channel <= 65535)
        nv = int(np.prod(dims))
        big = 65535 // 3 + 1
        J = np.zeros((nv * 3, big), dtype=np.complex128, order="F")
        with pytest.raises(L.VieError) as e:
            op.apply("N", J)
        assert e.value.code == L.VIE_ERR_ARG
    finally:
        op.close()
