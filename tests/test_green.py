"""Green-tensor set-up: near-field scheme (host) and far/medium quadrature (HIP).

Near field (marie_vie.green, integration by parts + Duffy boxes) is pinned by
  * the static self term of the PWC N operator.  The reference kernel writes
    the diagonal as -d_j^2 g - d_k^2 g (G_N.m:15-20); over the cube each
    d_j d_j of the Newtonian potential integrates to -h^3/3 (cubic symmetry
    of the Laplacian), so the entry is +2 h^3 / 3, off-diagonal zero;
  * the trace identity: summed over the diagonal the kernel is
    -2 laplacian(g) = 2 k0^2 g + 2 delta, so for every monomial pair the
    trace is 2 k0^2 int int T B g plus twice the overlap int T B on the self
    voxel, and without it on the voxels that touch it;
  * agreement with the reference's volume-volume Gauss rule (the oracle's
    vv_op, a restatement of VV_Nop.m / VV_Kop.m) at separated offsets where
    that rule converges.
The HIP quadrature kernel is compared with vv_op on a few offsets.
"""
import numpy as np
import pytest

from oracle import vie_oracle as vo


def test_near_self_term_static_limit():
    from marie_vie import green as gr
    h = 0.005
    b = gr._Boxes(h, 1e-9)
    vals = [gr.near_entry(b, (0, 0, 0), 0, 0, "N", c) / h ** 3 for c in range(6)]
    for c in (0, 3, 5):
        assert abs(vals[c] - 2.0 / 3.0) < 1e-10
    for c in (1, 2, 4):
        assert abs(vals[c]) < 1e-12


@pytest.mark.parametrize("m", [(0, 0, 0), (1, 0, 0), (0, 1, 1), (1, 1, 1)])
def test_near_trace_identity(m):
    from marie_vie import green as gr
    h, k0 = 0.004, 6.2456
    b = gr._Boxes(h, k0)
    gram = {0: h ** 3, 1: h ** 3 / 12, 2: h ** 3 / 12, 3: h ** 3 / 12}
    for l, lp in gr.PAIRS:
        tr = sum(gr.near_entry(b, m, l, lp, "N", c) for c in (0, 3, 5))
        td = [1 if gr._mono(l) == d else 0 for d in range(3)]
        bd = [1 if gr._mono(lp) == d else 0 for d in range(3)]
        vol = gr._integral(b, [gr.Axis(("V",), m[d], td[d], bd[d], h) for d in range(3)])
        want = 2 * k0 ** 2 * vol + (2 * gram[l] if (m == (0, 0, 0) and l == lp) else 0.0)
        assert abs(tr - want) <= 1e-9 * h ** 3, (m, l, lp, tr, want)


@pytest.mark.parametrize("op", ["N", "K"])
def test_near_scheme_matches_volume_rule_when_separated(op):
    from marie_vie import green as gr
    h, k0 = 0.004, 6.2456
    b = gr._Boxes(h, k0)
    m = (3, 1, 2)
    ref = vo.vv_op(m, h, k0, 7, op, True)
    for p, (l, lp) in enumerate(gr.PAIRS):
        for c in range(ref.shape[1]):
            got = gr.near_entry(b, m, l, lp, op, c)
            assert abs(got - ref[p, c]) <= 1e-8 * np.abs(ref).max()


def test_near_adjacent_offsets_converge():
    """Touching voxels (offset (1,0,0)): the scheme is stable under refinement
    of its own quadrature (12 vs 16 points per Duffy coordinate)."""
    from marie_vie import green as gr
    h, k0 = 0.004, 2.68
    b1 = gr._Boxes(h, k0, 8, 12)
    b2 = gr._Boxes(h, k0, 10, 16)
    for (l, lp) in gr.PAIRS[:5]:
        for c in range(6):
            a = gr.near_entry(b1, (1, 0, 0), l, lp, "N", c)
            r = gr.near_entry(b2, (1, 0, 0), l, lp, "N", c)
            assert abs(a - r) <= 1e-9 * max(abs(r), h ** 3)


@pytest.mark.gpu
@pytest.mark.parametrize("op", ["N", "K"])
@pytest.mark.parametrize("pwl", [False, True])
def test_gpu_quadrature_matches_vv_rule(op, pwl):
    import ctypes
    from marie_vie import VieOperator
    from marie_vie import _lib as L
    dims = (5, 4, 6)
    h, k0 = 0.004, 2.6827
    opr = VieOperator(dims, "PWL" if pwl else "PWC", h)
    pwx = 10 if pwl else 1
    nc = 6 if op == "N" else 3
    A = np.zeros(dims + (pwx, nc), dtype=np.complex128, order="F")
    nq = 4
    x, w = np.polynomial.legendre.leggauss(nq)
    rc = opr.lib.vie_green_quadrature(opr.h, L.VIE_OP_N if op == "N" else L.VIE_OP_K, k0, nq,
                                      ctypes.c_void_p(x.ctypes.data), ctypes.c_void_p(w.ctypes.data),
                                      dims[0], dims[1], dims[2], ctypes.c_void_p(A.ctypes.data), L.VIE_MEM_HOST)
    opr._check(rc)
    for m in [(1, 1, 1), (4, 0, 2), (2, 3, 5), (0, 1, 3)]:
        ref = vo.vv_op(m, h, k0, nq, op, pwl)
        got = A[m[0], m[1], m[2]]
        assert np.abs(got - ref).max() <= 1e-12 * np.abs(ref).max()
    opr.close()


@pytest.mark.parametrize("dims", [(1, 5, 4), (2, 9, 9), (3, 10, 12), (5, 40, 40), (6, 9, 9), (7, 20, 3),
                                  (12, 12, 12), (64, 64, 80), (96, 96, 112)])
@pytest.mark.parametrize("op", ["N", "K"])
def test_assembly_regimes_match_reference_loops(op, dims):
    """Which rule (far / medium of which order / near) the library applies at
    each offset equals what the loops of assembly_N.m / assembly_K.m write last
    inside the grid: N shrinks the medium order when some n <= 5, K only when
    some n <= 2 (so K keeps order 8 on a 3..5-voxel axis)."""
    from marie_vie import green
    rule, order = vo.assembly_rules(op, dims)
    got_rule = np.zeros(dims, dtype=int)
    got_order = np.full(dims, 4, dtype=int)
    nm, ext = green.medium_rule(op, dims)
    got_rule[:ext[0], :ext[1], :ext[2]] = 1
    got_order[:ext[0], :ext[1], :ext[2]] = nm
    for m in green.near_offsets(dims):
        got_rule[m] = 2
        got_order[m] = 0
    np.testing.assert_array_equal(got_rule, rule)
    np.testing.assert_array_equal(got_order, order)
    if op == "K" and min(dims) in (3, 4, 5):
        assert nm == 8


def test_green_tensor_save_load_round_trip(tmp_path):
    from marie_vie import green
    rng = np.random.default_rng(4)
    dims = (3, 4, 2)
    A = rng.standard_normal(dims + (10, 6)) + 1j * rng.standard_normal(dims + (10, 6))
    p = tmp_path / "N.npz"
    green.save(p, A, "N", dims, 0.004, 2.6827, True)
    B = green.load(p, "N", dims, 0.004, 2.6827, True)
    np.testing.assert_array_equal(A, B)
    for bad in [("K", dims, 0.004, 2.6827, True), ("N", (3, 4, 3), 0.004, 2.6827, True),
                ("N", dims, 0.005, 2.6827, True), ("N", dims, 0.004, 6.2456, True)]:
        with pytest.raises(ValueError):
            green.load(p, *bad)
    with pytest.raises(ValueError):
        green.save(p, A, "N", dims, 0.004, 2.6827, False)      # wrong pair count for PWC
