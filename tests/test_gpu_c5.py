"""Size-independent properties at the body-scale workload (C5: 128 x 128 x
160 grid of 4 mm voxels, full-body ellipsoid stack, PWL, 123 MHz, Green
tensor assembled on the device), where no CPU oracle solve is feasible:
batched (16 rhs) == single bitwise, <A x, y> = <x, A^H y> in FP64, FP32
within the FP32 tolerance of FP64, and a short batched mixed-precision
solve that lowers every residual; the same operator properties at the
torso workload (C4, PWC) and its support box against the full grid."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c5_scale_operator_properties():
    import torch
    from marie_vie import VieOperator, green, phantoms
    freq = 123e6
    ph = phantoms.body_stack()
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWL", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(freq)
    op.set_material(Mr, Mc, ph.idxS)
    op.use_torch_stream()
    n = op.packed_len()
    print(f"C5 dims={ph.dims} nS={len(ph.idxS)} unknowns/rhs={n}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(5005)
    X = torch.randn(16, n, dtype=torch.complex128, device="cuda", generator=g)
    Y = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    Yb = op.apply("VIE", X, precision="fp32")
    for k in (0, 11):
        one = op.apply("VIE", X[k].contiguous(), precision="fp32")
        assert torch.equal(Yb[k], one)
    Ax = op.apply("VIE", X[0].contiguous())
    AhY = op.apply("VIEH", Y)
    lhs = torch.vdot(Y, Ax)
    rhs = torch.vdot(AhY, X[0])
    assert abs(complex(lhs - rhs)) / abs(complex(lhs)) < 1e-12
    d32 = float(torch.linalg.vector_norm(Yb[0] - Ax) / torch.linalg.vector_norm(Ax))
    assert d32 < 2e-5
    x, info = op.gmres(X, tol=1e-2, restart=20, maxit=5, precision="mixed")
    for k in range(16):
        r = float(torch.linalg.vector_norm(X[k] - op.apply("VIE", x[k].contiguous())) / torch.linalg.vector_norm(X[k]))
        assert r < 1e-1, (k, r)
    print(f"C5 fp32 vs fp64 {d32:.2e}; mixed solve matvecs {info['nmatvec'][:4]}...", flush=True)
    op.close()


def test_c4_scale_operator_properties(monkeypatch):
    """The torso workload (C4: 96 x 96 x 112 grid of 3 mm voxels, layered
    shells, PWC, 298 MHz, Green tensor assembled on the device), whose support
    box runs the x pass at length 192 (split 12 x 16) and the FP32 axis passes
    at 160 and 224: 8 rhs in one launch == one at a time bitwise (FP64 and
    FP32), <A x, y> = <x, A^H y> in FP64, FP32 within the FP32 tolerance of
    FP64, and the support box equal to the full grid to rounding."""
    import torch
    from marie_vie import VieOperator, green, phantoms
    freq = 298e6
    ph = phantoms.layered_torso()
    const = phantoms.constants(freq)
    Mr, Mc = ph.material(freq)

    def make():
        op = VieOperator(ph.dims, "PWC", ph.res)
        op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
        op.set_material(Mr, Mc, ph.idxS)
        op.use_torch_stream()
        return op

    op = make()
    box = op.box()
    assert box is not None and box["fft"][0] == 192, box
    n = op.packed_len()
    print(f"C4 dims={ph.dims} nS={len(ph.idxS)} unknowns/rhs={n} box={box}", flush=True)
    g = torch.Generator(device="cuda").manual_seed(4004)
    X = torch.randn(8, n, dtype=torch.complex128, device="cuda", generator=g)
    Y = torch.randn(n, dtype=torch.complex128, device="cuda", generator=g)
    for prec in ("fp64", "fp32"):
        Yb = op.apply("VIE", X, precision=prec)
        for k in (0, 5):
            assert torch.equal(Yb[k], op.apply("VIE", X[k].contiguous(), precision=prec)), (prec, k)
    Ax = op.apply("VIE", X[0].contiguous())
    AhY = op.apply("VIEH", Y)
    lhs = torch.vdot(Y, Ax)
    rhs = torch.vdot(AhY, X[0])
    assert abs(complex(lhs - rhs)) / abs(complex(lhs)) < 1e-12
    A32 = op.apply("VIE", X[0].contiguous(), precision="fp32")
    d32 = float(torch.linalg.vector_norm(A32 - Ax) / torch.linalg.vector_norm(Ax))
    assert d32 < 2e-5
    op.close()
    monkeypatch.setenv("VIE_BOX", "0")
    full = make()
    assert full.box() is None
    Af = full.apply("VIE", X[0].contiguous())
    dfull = float((Af - Ax).abs().max() / Ax.abs().max())
    assert dfull < 1e-13
    full.close()
    print(f"C4 fp32 vs fp64 {d32:.2e}; box vs full grid {dfull:.2e}", flush=True)
