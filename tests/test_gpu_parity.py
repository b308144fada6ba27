"""GPU parity: the HIP operator (through the C ABI) against the CPU oracle.

Tolerance: the plain FP64 form must match to 1e-12 relative (max-abs error
over max-abs value); the FP32 form to 2e-5.  Inputs are seeded; the real-space
kernels are random complex tensors with the reference's parity structure
(tests/bruteforce.py), and the oracle itself is pinned by brute-force
Toeplitz products in tests/test_oracle.py.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo
import bruteforce as bf

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 2e-5


def _rel(a, b):
    return np.abs(a - b).max() / np.abs(b).max()


def _op(dims, pwl, res=0.005):
    from marie_vie import VieOperator
    return VieOperator(dims, "PWL" if pwl else "PWC", res)


def _field(rng, dims, ql, nrhs=None):
    shape = dims + (ql,) + ((nrhs,) if nrhs else ())
    return np.asfortranarray(rng.standard_normal(shape) + 1j * rng.standard_normal(shape))


GRIDS = [(3, 4, 5), (2, 3, 2), (8, 6, 5), (11, 4, 7), (16, 12, 10)]


@pytest.mark.parametrize("dims", GRIDS)
@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("opname", ["N", "K"])
def test_dense_spectrum_ops_fp64(dims, pwl, opname):
    print(f"case dims={dims} pwl={pwl} op={opname}", flush=True)
    rng = np.random.default_rng(100 + sum(dims) + 7 * pwl)
    A = bf.random_kernel(rng, dims, opname, pwl)
    T = vo.circulant_dense(A, pwl)
    op = _op(dims, pwl)
    op.set_green_dense(opname, A)
    ql = 12 if pwl else 3
    J = _field(rng, dims, ql)
    ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
    got = op.apply(opname, J)
    assert _rel(got, ref_apply(J, T, pwl)) < TOL64
    gotH = op.apply(opname + "H", J)
    assert _rel(gotH, ref_apply(J, T, pwl, adjoint=True)) < TOL64
    got32 = op.apply(opname, J, precision="fp32")
    assert _rel(got32, ref_apply(J, T, pwl)) < TOL32
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
@pytest.mark.parametrize("opname", ["N", "K"])
def test_tucker_tstruct_ops_fp64(pwl, opname):
    """Feed the reference T-struct (HOSVD at tol 1e-7) to the library; the
    oracle re-expands the same T-struct on every matvec like the reference."""
    dims = (6, 8, 5)
    rng = np.random.default_rng(7 + pwl)
    A = bf.random_kernel(rng, dims, opname, pwl)
    T = (vo.fft_circ_tucker_pwl if pwl else vo.fft_circ_tucker_pwc)(A, 1e-7)
    op = _op(dims, pwl)
    op.set_green_tucker(opname, T)
    ql = 12 if pwl else 3
    J = _field(rng, dims, ql)
    ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
    assert _rel(op.apply(opname, J), ref_apply(J, T, pwl)) < TOL64
    assert _rel(op.apply(opname + "H", J), ref_apply(J, T, pwl, adjoint=True)) < TOL64
    op.close()


def _material(rng, dims, frac=0.5, freq=298e6):
    nv = int(np.prod(dims))
    eps = 1 + 60 * rng.random(dims)
    sig = rng.random(dims)
    ce = vo.em_constants(freq)["ce"]
    mat = vo.material(eps, sig, ce)
    idxS = np.sort(rng.choice(nv, size=max(1, int(frac * nv)), replace=False))
    return mat, idxS


@pytest.mark.parametrize("dims", [(3, 4, 5), (10, 9, 8)])
@pytest.mark.parametrize("pwl", [False, True])
def test_vie_operator_and_adjoint_fp64(dims, pwl):
    print(f"case dims={dims} pwl={pwl}", flush=True)
    rng = np.random.default_rng(31 + pwl + dims[0])
    A = bf.random_kernel(rng, dims, "N", pwl)
    T = vo.circulant_dense(A, pwl)
    res = 0.005
    op = _op(dims, pwl, res)
    op.set_green_dense("N", A)
    mat, idxS = _material(rng, dims)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    ql = 12 if pwl else 3
    isol = vo.idx_solution(idxS, int(np.prod(dims)), ql)
    x = rng.standard_normal(len(isol)) + 1j * rng.standard_normal(len(isol))
    got = op.apply("VIE", x)
    ref = vo.mvp_vie(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    assert _rel(got, ref) < TOL64
    gotH = op.apply("VIEH", x)
    refH = vo.mvp_vie_adjoint(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    assert _rel(gotH, refH) < TOL64
    got32 = op.apply("VIE", x, precision="fp32")
    assert _rel(got32, ref) < TOL32
    op.close()


def test_batched_equals_single():
    dims = (8, 7, 6)
    rng = np.random.default_rng(5)
    A = bf.random_kernel(rng, dims, "N", True)
    op = _op(dims, True)
    op.set_green_dense("N", A)
    J = _field(rng, dims, 12, nrhs=3)
    got = op.apply("N", J)
    for k in range(3):
        one = op.apply("N", np.asfortranarray(J[..., k]))
        np.testing.assert_array_equal(got[..., k], one)
    mat, idxS = _material(rng, dims)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    x = np.asfortranarray(rng.standard_normal((12 * len(idxS), 4)) + 1j * rng.standard_normal((12 * len(idxS), 4)))
    gb = op.apply("VIE", x)
    for k in range(4):
        np.testing.assert_array_equal(gb[:, k], op.apply("VIE", np.ascontiguousarray(x[:, k])))
    op.close()


def test_c1_grid_pwc_n_fp64():
    """C1 shape (32^3 PWC): parity against the oracle on a random kernel."""
    dims = (32, 32, 32)
    rng = np.random.default_rng(1001)
    A = bf.random_kernel(rng, dims, "N", False)
    T = vo.circulant_dense(A, False)
    op = _op(dims, False)
    op.set_green_dense("N", A)
    J = _field(rng, dims, 3)
    assert _rel(op.apply("N", J), vo.mvp_N(J, T, False)) < TOL64
    op.close()


def test_adjoint_identity_at_c2_size():
    """Size-independent check at the 64x64x80 PWL grid: <A x, y> = <x, A^H y>."""
    dims = (64, 64, 80)
    rng = np.random.default_rng(1002)
    A = bf.random_kernel(rng, dims, "N", True)
    op = _op(dims, True, 0.004)
    op.set_green_dense("N", A)
    mat, idxS = _material(rng, dims, frac=0.17)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    n = op.packed_len()
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    lhs = np.vdot(y, op.apply("VIE", x))
    rhs = np.vdot(op.apply("VIEH", y), x)
    assert abs(lhs - rhs) / abs(lhs) < 1e-12
    # linearity
    a = 0.3 - 1.7j
    z = op.apply("VIE", a * x + y)
    np.testing.assert_allclose(z, a * op.apply("VIE", x) + op.apply("VIE", y), rtol=0,
                               atol=1e-12 * np.abs(z).max())
    op.close()


def test_gmres_matches_oracle_iterations():
    dims = (6, 5, 7)
    rng = np.random.default_rng(77)
    A = bf.random_kernel(rng, dims, "N", False)
    res = 0.005
    mat, idxS = _material(rng, dims, frac=0.6)
    isol = vo.idx_solution(idxS, int(np.prod(dims)), 3)
    b = rng.standard_normal(len(isol)) + 1j * rng.standard_normal(len(isol))
    # scale the kernel so that the compact part has norm ~0.6 (GMRES converges
    # in a few restart cycles, which exercises the restart bookkeeping)
    T0 = vo.circulant_dense(A, False)
    d = b - vo.mvp_vie(b, mat["Mr"], mat["Mc"], isol, dims, res, T0, False)
    A = A * (0.6 * np.linalg.norm(b) / np.linalg.norm(d))
    T = vo.circulant_dense(A, False)
    op = _op(dims, False, res)
    op.set_green_dense("N", A)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    afun = lambda v: vo.mvp_vie(v, mat["Mr"], mat["Mc"], isol, dims, res, T, False)
    xr, itr, rvr = vo.gmres_vie(afun, b, restart=8, tol=1e-10, maxit=50)
    x, info = op.gmres(b, tol=1e-10, restart=8, maxit=50)
    assert info["iters"][0] == list(itr)
    assert len(info["resvec"][0]) == len(rvr)
    np.testing.assert_allclose(info["resvec"][0], rvr, rtol=1e-6, atol=1e-14)
    assert _rel(x, xr) < 1e-9
    op.close()


def test_gmres_batched_matches_single_solves():
    """Three rhs in one batch whose cycles end at different steps (the active
    set shrinks between cycles): FP64 reproduces the oracle per rhs, and the
    mixed-precision refinement reaches its tolerance on every rhs."""
    dims = (6, 5, 7)
    rng = np.random.default_rng(78)
    A = bf.random_kernel(rng, dims, "N", False)
    res = 0.005
    mat, idxS = _material(rng, dims, frac=0.6)
    isol = vo.idx_solution(idxS, int(np.prod(dims)), 3)
    n = len(isol)
    T0 = vo.circulant_dense(A, False)
    b0 = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    d = b0 - vo.mvp_vie(b0, mat["Mr"], mat["Mc"], isol, dims, res, T0, False)
    A = A * (0.6 * np.linalg.norm(b0) / np.linalg.norm(d))
    T = vo.circulant_dense(A, False)
    afun = lambda v: vo.mvp_vie(v, mat["Mr"], mat["Mc"], isol, dims, res, T, False)
    # rhs 1: mostly a fixed point of the compact part (few steps); rhs 2: random
    B = np.stack([b0, afun(b0) + 1e-3 * b0, rng.standard_normal(n) + 1j * rng.standard_normal(n)])
    op = _op(dims, False, res)
    op.set_green_dense("N", A)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    x, info = op.gmres(B.T, tol=1e-10, restart=8, maxit=50)       # rhs as columns
    iters = []
    for i in range(3):
        xr, itr, rvr = vo.gmres_vie(afun, B[i], restart=8, tol=1e-10, maxit=50)
        iters.append(list(itr))
        print(f"batched gmres rhs {i}: oracle iters {list(itr)} device {info['iters'][i]}", flush=True)
        assert info["iters"][i] == list(itr)
        np.testing.assert_allclose(info["resvec"][i], rvr, rtol=1e-6, atol=1e-14)
        assert _rel(x[:, i], xr) < 1e-9
    assert len({tuple(t) for t in iters}) > 1          # the rhs finish at different steps
    xm, infm = op.gmres(B.T, tol=1e-8, restart=8, maxit=50, precision="mixed")
    for i in range(3):
        r = B[i] - afun(xm[:, i])
        assert np.linalg.norm(r) <= 1e-8 * np.linalg.norm(B[i])
        assert _rel(xm[:, i], x[:, i]) < 1e-6
    op.close()


@pytest.mark.parametrize("dims", [(32, 64, 80), (11, 12, 10)])
def test_support_pruning_bitwise_equal(dims, monkeypatch):
    """Packed operators skip the z lines without scatterer voxels (zero input,
    outputs never gathered); VIE_PRUNE=0 runs every line.  Bitwise equal, and
    the pruned result matches the oracle."""
    print(f"case pruning dims={dims}", flush=True)
    monkeypatch.setenv("VIE_BOX", "0")   # same grid in both runs (the box test covers the box)
    monkeypatch.setenv("VIE_AXIS_FFT", "ct")   # both contexts on one FFT schedule (it rounds differently)
    rng = np.random.default_rng(616)
    pwl = dims[0] == 32
    A = bf.random_kernel(rng, dims, "N", pwl)
    ql = 12 if pwl else 3
    # a compact ellipsoidal support: most z lines hold no scatterer
    c = [(np.arange(n) - (n - 1) / 2) / (n / 2) for n in dims]
    X, Y, Z = np.meshgrid(*c, indexing="ij")
    inside = (X / 0.6) ** 2 + (Y / 0.5) ** 2 + (Z / 0.7) ** 2 <= 1.0
    idxS = np.nonzero(inside.reshape(-1, order="F"))[0]
    nv = int(np.prod(dims))
    eps = 1 + 60 * rng.random(dims)
    mat = vo.material(eps, rng.random(dims), vo.em_constants(298e6)["ce"])
    n = ql * len(idxS)
    x = np.asfortranarray(rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3)))
    outs = {}
    for prune in ("0", "1"):
        monkeypatch.setenv("VIE_PRUNE", prune)
        op = _op(dims, pwl)
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        outs[prune] = [op.apply(o, v, precision=p) for o in ("VIE", "VIEH") for p in ("fp64", "fp32")
                       for v in (np.ascontiguousarray(x[:, 0]), x)]
        op.close()
    for a, b in zip(outs["1"], outs["0"]):
        np.testing.assert_array_equal(a, b)
    isol = vo.idx_solution(idxS, nv, ql)
    ref = vo.mvp_vie(x[:, 0], mat["Mr"], mat["Mc"], isol, dims, 0.005, vo.circulant_dense(A, pwl), pwl)
    assert _rel(outs["1"][0], ref) < TOL64


@pytest.mark.parametrize("dims", [(40, 48, 56), (32, 24, 20)])
def test_x_stage_bitwise_equal(dims, monkeypatch):
    """The x pass with the block's octant spectrum column staged in LDS (PWL N
    at x lengths 80 and 64, FP64 and FP32) equals the plain kernel
    (VIE_XSTAGE=0) bitwise: N, NH, VIE, VIEH, one rhs and three (an odd batch
    puts the two columns of some blocks on different octant columns, which
    read their spectrum from memory); the staged N matches the oracle."""
    print(f"case x stage dims={dims}", flush=True)
    monkeypatch.setenv("VIE_BOX", "0")         # the packed operators on the grid's own lengths
    monkeypatch.setenv("VIE_AXIS_FFT", "ct")   # both contexts on one axis-pass schedule
    rng = np.random.default_rng(717 + dims[0])
    A = bf.random_kernel(rng, dims, "N", True)
    mat, idxS = _material(rng, dims, frac=0.4)
    nv = int(np.prod(dims))
    J = _field(rng, dims, 12, 3)
    n = 12 * len(idxS)
    x = np.asfortranarray(rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3)))
    outs = {}
    for stage in ("0", "1"):
        monkeypatch.setenv("VIE_XSTAGE", stage)
        op = _op(dims, True)
        assert op.fft_dims[0] == 2 * dims[0]
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        res = []
        for p in ("fp64", "fp32"):
            for o in ("N", "NH"):
                res += [op.apply(o, np.asfortranarray(J[..., 0]), precision=p), op.apply(o, J, precision=p)]
            for o in ("VIE", "VIEH"):
                res += [op.apply(o, np.ascontiguousarray(x[:, 0]), precision=p), op.apply(o, x, precision=p)]
        outs[stage] = res
        op.close()
    for a, b in zip(outs["1"], outs["0"]):
        np.testing.assert_array_equal(a, b)
    ref = vo.mvp_N(np.asfortranarray(J[..., 0]), vo.circulant_dense(A, True), True)
    assert _rel(outs["1"][0], ref) < TOL64
    assert _rel(outs["1"][8], ref) < TOL32
    isol = vo.idx_solution(idxS, nv, 12)
    refv = vo.mvp_vie(x[:, 0], mat["Mr"], mat["Mc"], isol, dims, 0.005, vo.circulant_dense(A, True), True)
    assert _rel(outs["1"][4], refv) < TOL64


@pytest.mark.parametrize("support", ["slab", "edges", "single"])
@pytest.mark.parametrize("dims", [(32, 12, 10), (11, 12, 10)])
def test_support_pruning_x_range(dims, support, monkeypatch):
    """The y and x passes of the packed operators touch only the scatterer's
    x range: a two-plane slab inside one tile width, voxels on both x edges
    of the grid, one voxel.  Pruned == VIE_PRUNE=0 bitwise (FP64 and FP32,
    VIE and VIE^H, one rhs and a batch), and FP64 matches the oracle
    (n1 = 32: compile-time x pass at P = 64; 11 x 12 x 10: generic kernels)."""
    print(f"case x-range pruning dims={dims} support={support}", flush=True)
    monkeypatch.setenv("VIE_AXIS_FFT", "rf")   # both contexts on one FFT schedule
    monkeypatch.setenv("VIE_BOX", "0")
    rng = np.random.default_rng(617)
    pwl = dims[0] == 32
    A = bf.random_kernel(rng, dims, "N", pwl)
    ql = 12 if pwl else 3
    n1, n2, n3 = dims
    inside = np.zeros(dims, bool)
    if support == "slab":
        inside[5:7, 2:n2 - 3, 1:n3 - 2] = True
    elif support == "edges":
        inside[0, 3, 4] = inside[0, 7, 2] = inside[n1 - 1, 5, 5] = inside[n1 - 1, 6, 8] = True
    else:
        inside[n1 // 2, n2 // 2, n3 // 2] = True
    idxS = np.nonzero(inside.reshape(-1, order="F"))[0]
    nv = int(np.prod(dims))
    eps = 1 + 60 * rng.random(dims)
    mat = vo.material(eps, rng.random(dims), vo.em_constants(298e6)["ce"])
    n = ql * len(idxS)
    x = np.asfortranarray(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))
    outs = {}
    for prune in ("0", "1"):
        monkeypatch.setenv("VIE_PRUNE", prune)
        op = _op(dims, pwl)
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        outs[prune] = [op.apply(o, v, precision=p) for o in ("VIE", "VIEH") for p in ("fp64", "fp32")
                       for v in (np.ascontiguousarray(x[:, 0]), x)]
        op.close()
    for a, b in zip(outs["1"], outs["0"]):
        np.testing.assert_array_equal(a, b)
    isol = vo.idx_solution(idxS, nv, ql)
    T = vo.circulant_dense(A, pwl)
    ref = vo.mvp_vie(x[:, 0], mat["Mr"], mat["Mc"], isol, dims, 0.005, T, pwl)
    assert _rel(outs["1"][0], ref) < TOL64


@pytest.mark.parametrize("dims,lo,hi,source", [((64, 64, 80), (12, 8, 13), (51, 55, 66), "dense"),
                                                ((20, 18, 24), (0, 3, 4), (19, 10, 15), "dense"),
                                                ((20, 18, 24), (0, 3, 4), (19, 10, 15), "tucker")])
def test_support_box_matches_full_grid(dims, lo, hi, source, monkeypatch):
    """Packed operators on the scatterer's bounding box (kernel offsets < w_d,
    FFT lengths >= 2 w_d) against the same operators on the full grid
    (VIE_BOX=0): VIE and VIE^H in FP64 to 1e-13 and FP32 to 2e-6 of each
    other, batches as single rhs, GMRES with the same [p outit] and resvec,
    and FP64 against the oracle (64^2 x 80 PWL: box 40 x 48 x 54 on FFT
    lengths 80 x 96 x 112; 20 x 18 x 24 PWC: box 20 x 8 x 12 on 64 x 16 x 24,
    compile-time x pass with generic y and z lengths), the kernel given dense
    or as the reference's Tucker T-struct (rebuilt in real space per term)."""
    print(f"case support box dims={dims} box {lo}-{hi} source={source}", flush=True)
    rng = np.random.default_rng(618)
    pwl = dims[0] == 64
    A = bf.random_kernel(rng, dims, "N", pwl)
    ql = 12 if pwl else 3
    if source == "tucker":
        T = (vo.fft_circ_tucker_pwl if pwl else vo.fft_circ_tucker_pwc)(A, 1e-9)
    inside = np.zeros(dims, bool)
    sl = tuple(slice(a, b + 1) for a, b in zip(lo, hi))
    inside[sl] = rng.random(inside[sl].shape) < 0.7
    inside[lo] = inside[hi] = True   # two corners fix the bounding box
    idxS = np.nonzero(inside.reshape(-1, order="F"))[0]
    eps = 1 + 30 * rng.random(dims)
    mat = vo.material(eps, rng.random(dims), vo.em_constants(298e6)["ce"])
    n = ql * len(idxS)
    x = np.asfortranarray(rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2)))
    b = np.ascontiguousarray(x[:, 0])
    res = {}
    for box in ("0", "1"):
        monkeypatch.setenv("VIE_BOX", box)
        op = _op(dims, pwl)
        if source == "tucker":
            op.set_green_tucker("N", T)
        else:
            op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        info_box = op.box()
        if box == "1":
            w = tuple(h - l + 1 for l, h in zip(lo, hi))
            assert info_box is not None and info_box["lo"] == lo and info_box["extent"] == w, info_box
            print(f"box {info_box}", flush=True)
        else:
            assert info_box is None
        ops = [op.apply(o, v, precision=p) for o in ("VIE", "VIEH") for p in ("fp64", "fp32") for v in (b, x)]
        xs, info = op.gmres(b, tol=1e-8, restart=20, maxit=10)
        res[box] = (ops, xs, info)
        op.close()
    for k, (p_box, p_full) in enumerate(zip(res["1"][0], res["0"][0])):
        tol = 1e-13 if (k // 2) % 2 == 0 else 2e-6
        assert _rel(p_box, p_full) < tol, (k, _rel(p_box, p_full))
    assert np.array_equal(res["1"][0][1][:, 0], res["1"][0][0])          # batch column 0 == single rhs
    assert res["1"][2]["iters"][0] == res["0"][2]["iters"][0]
    np.testing.assert_allclose(res["1"][2]["resvec"][0], res["0"][2]["resvec"][0], rtol=1e-8, atol=1e-15)
    assert _rel(res["1"][1], res["0"][1]) < 1e-10
    if not pwl:   # (the full-grid operator is itself pinned to the oracle at every size)
        isol = vo.idx_solution(idxS, int(np.prod(dims)), ql)
        Tref = T if source == "tucker" else vo.circulant_dense(A, pwl)
        ref = vo.mvp_vie(b, mat["Mr"], mat["Mc"], isol, dims, 0.005, Tref, pwl)
        assert _rel(res["1"][0][0], ref) < TOL64


@pytest.mark.parametrize("precision", ["fp64", "mixed"])
def test_gmres_graphs_match_eager(precision, monkeypatch):
    """The Arnoldi steps run as captured graph blocks (replayed across cycles
    and solves); VIE_GRAPHS=0 launches the same kernels eagerly.  Every replay,
    with eager work of the program in between, equals the eager solve bitwise."""
    dims = (6, 5, 7)
    rng = np.random.default_rng(79)
    A = bf.random_kernel(rng, dims, "N", False)
    mat, idxS = _material(rng, dims, frac=0.6)
    n = 3 * len(idxS)
    B = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    isol = vo.idx_solution(idxS, int(np.prod(dims)), 3)
    d = B[:, 0] - vo.mvp_vie(B[:, 0], mat["Mr"], mat["Mc"], isol, dims, 0.005, vo.circulant_dense(A, False), False)
    A = A * (0.6 * np.linalg.norm(B[:, 0]) / np.linalg.norm(d))      # compact part of norm ~0.6
    res = {}
    monkeypatch.setenv("VIE_AXIS_FFT", "ct")   # both contexts on one FFT schedule
    for graphs in ("0", "1"):
        monkeypatch.setenv("VIE_GRAPHS", graphs)
        op = _op(dims, False)
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        runs = []
        for rep in range(3):
            x, info = op.gmres(np.asfortranarray(B), tol=1e-10, restart=8, maxit=50, precision=precision)
            runs.append((x, info))
            op.apply("VIE", np.ascontiguousarray(B[:, 0]))          # eager work between replays
        op.close()
        res[graphs] = runs
    x0, i0 = res["0"][0]
    assert min(i0["nmatvec"]) > 16                                   # several restart cycles of 8
    for graphs in ("0", "1"):
        for x, info in res[graphs]:
            np.testing.assert_array_equal(x, x0)
            assert info["iters"] == i0["iters"] and info["nmatvec"] == i0["nmatvec"]
            for a, b in zip(info["resvec"], i0["resvec"]):
                np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("dims,pwl", [((31, 6, 5), True), ((31, 6, 5), False), ((95, 4, 3), False)])
def test_padded_fft_lengths(dims, pwl):
    """n1 = 31 and 95 have no 7-smooth 2 n1, so the x axis runs at P = 64 / 192
    (compile-time kernels) with P > 2 n1: the inverse keeps only n1 outputs."""
    print(f"case padded dims={dims} pwl={pwl}", flush=True)
    rng = np.random.default_rng(90 + dims[0] + pwl)
    op = _op(dims, pwl)
    P1 = op.fft_dims[0]
    assert P1 > 2 * dims[0]
    ql = 12 if pwl else 3
    for opname in ("N", "K"):
        A = bf.random_kernel(rng, dims, opname, pwl)
        T = vo.circulant_dense(A, pwl)
        op.set_green_dense(opname, A)
        J = _field(rng, dims, ql)
        ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
        assert _rel(op.apply(opname, J), ref_apply(J, T, pwl)) < TOL64
        assert _rel(op.apply(opname + "H", J), ref_apply(J, T, pwl, adjoint=True)) < TOL64
        assert _rel(op.apply(opname, J, precision="fp32"), ref_apply(J, T, pwl)) < TOL32
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
def test_x_pass_split_12x16(pwl):
    """x length 192 (n1 = 96) runs the x pass split 12 x 16 (PWC five
    columns per block, PWL one): N, K, their adjoints and FP32 against the
    oracle; the packed VIE operator against the oracle, and on 3 rhs in one
    launch equal to one rhs at a time bitwise (FP64 and FP32)."""
    dims = (96, 4, 3) if pwl else (96, 5, 4)
    ql = 12 if pwl else 3
    print(f"case x split 12x16 dims={dims} pwl={pwl}", flush=True)
    rng = np.random.default_rng(4192 + pwl)
    op = _op(dims, pwl)
    assert op.fft_dims[0] == 192
    TN = None
    for opname in ("N", "K"):
        A = bf.random_kernel(rng, dims, opname, pwl)
        T = vo.circulant_dense(A, pwl)
        if opname == "N":
            TN = T
        op.set_green_dense(opname, A)
        J = _field(rng, dims, ql)
        ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
        assert _rel(op.apply(opname, J), ref_apply(J, T, pwl)) < TOL64
        assert _rel(op.apply(opname + "H", J), ref_apply(J, T, pwl, adjoint=True)) < TOL64
        assert _rel(op.apply(opname, J, precision="fp32"), ref_apply(J, T, pwl)) < TOL32
    mat, idxS = _material(rng, dims)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    isol = vo.idx_solution(idxS, int(np.prod(dims)), ql)
    x = np.asfortranarray(rng.standard_normal((len(isol), 3)) + 1j * rng.standard_normal((len(isol), 3)))
    ref = vo.mvp_vie(np.ascontiguousarray(x[:, 0]), mat["Mr"], mat["Mc"], isol, dims, 0.005, TN, pwl)
    assert _rel(op.apply("VIE", np.ascontiguousarray(x[:, 0])), ref) < TOL64
    assert _rel(op.apply("VIE", np.ascontiguousarray(x[:, 0]), precision="fp32"), ref) < TOL32
    for prec in ("fp64", "fp32"):
        gb = op.apply("VIE", x, precision=prec)
        for k in range(3):
            np.testing.assert_array_equal(gb[:, k], op.apply("VIE", np.ascontiguousarray(x[:, k]), precision=prec))
    op.close()


@pytest.mark.parametrize("dims", [(32, 64, 80), (11, 12, 10)])
def test_tile_widths_bitwise_equal(dims, monkeypatch):
    """The axis-pass tile width (fixed by VIE_AXIS_TX, or autotuned) only
    changes which block computes which pencils: for each FFT schedule of the
    axis passes (VIE_AXIS_FFT=ct: Stockham stages, largest radix first; rf:
    two-step register split; cto: Stockham stages, odd radices first) results
    are bitwise equal for every width.  The schedules agree to rounding, and
    the tuner's own choice of schedule and width matches the oracle."""
    print(f"case tiles dims={dims}", flush=True)
    rng = np.random.default_rng(515)
    A = bf.random_kernel(rng, dims, "N", True)
    mat, idxS = _material(rng, dims, frac=0.3)
    n = 12 * len(idxS)
    x = rng.standard_normal((n, 2)) + 1j * rng.standard_normal((n, 2))

    def run():
        op = _op(dims, True)
        op.set_green_dense("N", A)
        op.set_material(mat["Mr"], mat["Mc"], idxS)
        tiles = (op.autotune("VIE", "fp32", 2), op.autotune("VIE", "fp64", 1))
        sched = (op.schedule("fp32"), op.schedule("fp64"))
        out = [op.apply("VIE", np.ascontiguousarray(x[:, 0])), op.apply("VIE", np.asfortranarray(x)),
               op.apply("VIE", np.asfortranarray(x), precision="fp32")]
        op.close()
        return out, tiles, sched

    outs = {}
    for fam in ("ct", "rf", "cto"):
        monkeypatch.setenv("VIE_AXIS_FFT", fam)
        for tx in ("", "4", "8", "16"):
            monkeypatch.setenv("VIE_AXIS_TX", tx)
            outs[fam, tx], tiles, sched = run()
            if tx == "":
                assert all(t in (4, 8, 16) for t in tiles[0]) and all(t in (4, 8) for t in tiles[1]), tiles
            assert sched == ([fam] * 4, [fam] * 4), sched
        for tx in ("4", "8", "16"):
            for a, b in zip(outs[fam, tx], outs[fam, ""]):
                np.testing.assert_array_equal(a, b)
    for other in ("rf", "cto"):
        for k, (a, b) in enumerate(zip(outs["ct", ""], outs[other, ""])):
            assert _rel(a, b) < (1e-13 if k < 2 else 1e-5), (other, k, _rel(a, b))
    monkeypatch.delenv("VIE_AXIS_FFT")
    monkeypatch.setenv("VIE_AXIS_TX", "")
    tuned, tiles, sched = run()
    assert all(s in ("ct", "rf", "cto") for s in sched[0] + sched[1]), sched
    print(f"tuned schedules fp32 {sched[0]} fp64 {sched[1]}, widths {tiles}", flush=True)
    isol = vo.idx_solution(idxS, int(np.prod(dims)), 12)
    ref = vo.mvp_vie(np.ascontiguousarray(x[:, 0]), mat["Mr"], mat["Mc"], isol, dims, 0.005,
                     vo.circulant_dense(A, True), True)
    assert _rel(tuned[0], ref) < TOL64
    assert _rel(tuned[2][:, 0], ref) < TOL32
    # one schedule per precision and pass: the batched application equals the single one bitwise
    np.testing.assert_array_equal(tuned[1][:, 0], tuned[0])


@pytest.mark.parametrize("generic", [False, True])
def test_specialised_lengths_pwl(generic, monkeypatch):
    """Grid 32x64x80 PWL: FFT lengths 64/128/160 run the compile-time kernels
    (generic=False) or, with VIE_GENERIC_KERNELS=1, the generic ones."""
    if generic:
        monkeypatch.setenv("VIE_GENERIC_KERNELS", "1")
    dims = (32, 64, 80)
    print(f"case specialised dims={dims} generic={generic}", flush=True)
    rng = np.random.default_rng(4242)
    res = 0.004
    op = _op(dims, True, res)
    for opname in ("N", "K"):
        A = bf.random_kernel(rng, dims, opname, True)
        T = vo.circulant_dense(A, True)
        op.set_green_dense(opname, A)
        J = _field(rng, dims, 12)
        ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
        assert _rel(op.apply(opname, J), ref_apply(J, T, True)) < TOL64
        assert _rel(op.apply(opname + "H", J), ref_apply(J, T, True, adjoint=True)) < TOL64
        assert _rel(op.apply(opname, J, precision="fp32"), ref_apply(J, T, True)) < TOL32
    op.close()


@pytest.mark.parametrize("sched", ["", "cto"])
@pytest.mark.parametrize("dims,pwl", [((40, 48, 56), True), ((56, 40, 48), False), ((48, 56, 40), True)])
def test_specialised_lengths_80_96_112(dims, pwl, sched, monkeypatch):
    """FFT lengths 80, 96 and 112 (those of the C2 support box) on each axis,
    with the x pass's split A x B = 10 x 8, 12 x 8, 14 x 8 (B not dividing A):
    N, K, their adjoints and FP32 against the oracle, with the tuned axis-pass
    schedules and with the Stockham stages in the odd-radix-first order
    (112 = 7 8 2, 96 = 3 8 4, 80 = 5 8 2)."""
    if sched:
        monkeypatch.setenv("VIE_AXIS_FFT", sched)
    print(f"case lengths 80/96/112 dims={dims} pwl={pwl} sched={sched or 'tuned'}", flush=True)
    rng = np.random.default_rng(4250 + dims[0])
    op = _op(dims, pwl, 0.004)
    ql = 12 if pwl else 3
    assert op.fft_dims == tuple(2 * n for n in dims)
    for opname in ("N", "K"):
        A = bf.random_kernel(rng, dims, opname, pwl)
        T = vo.circulant_dense(A, pwl)
        op.set_green_dense(opname, A)
        J = _field(rng, dims, ql)
        ref_apply = vo.mvp_N if opname == "N" else vo.mvp_K
        assert _rel(op.apply(opname, J), ref_apply(J, T, pwl)) < TOL64
        assert _rel(op.apply(opname + "H", J), ref_apply(J, T, pwl, adjoint=True)) < TOL64
        assert _rel(op.apply(opname, J, precision="fp32"), ref_apply(J, T, pwl)) < TOL32
    op.close()


@pytest.mark.parametrize("pwl", [False, True])
def test_specialised_lengths_vie(pwl):
    dims = (64, 32, 40) if pwl else (64, 64, 32)
    print(f"case specialised vie dims={dims} pwl={pwl}", flush=True)
    rng = np.random.default_rng(4243 + pwl)
    res = 0.004
    A = bf.random_kernel(rng, dims, "N", pwl)
    T = vo.circulant_dense(A, pwl)
    op = _op(dims, pwl, res)
    op.set_green_dense("N", A)
    mat, idxS = _material(rng, dims, frac=0.3)
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    ql = 12 if pwl else 3
    isol = vo.idx_solution(idxS, int(np.prod(dims)), ql)
    x = rng.standard_normal(len(isol)) + 1j * rng.standard_normal(len(isol))
    ref = vo.mvp_vie(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    assert _rel(op.apply("VIE", x), ref) < TOL64
    refH = vo.mvp_vie_adjoint(x, mat["Mr"], mat["Mc"], isol, dims, res, T, pwl)
    assert _rel(op.apply("VIEH", x), refH) < TOL64
    assert _rel(op.apply("VIE", x, precision="fp32"), ref) < TOL32
    op.close()
