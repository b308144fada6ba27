"""vie_apply_noalloc: the allocation-free variant of vie_apply.

Called through the C ABI on device tensors, it must give bitwise the results
of vie_apply (same kernels, workspaces in the library's fixed scratch array
instead of context buffers) on two calls in a row, the second on a smaller
input, while any device allocation by the library is made to fail
(VIE_FAULT_INJECT_OOM=1 makes the library's allocator refuse) and the
device's free memory stays unchanged.  A workspace beyond the scratch size is
refused with VIE_ERR_ARG before anything is launched.
"""
import ctypes

import numpy as np
import pytest

import bruteforce as bf
from oracle import vie_oracle as vo

pytestmark = pytest.mark.gpu


def _case(dims, pwl, seed):
    from marie_vie import VieOperator
    rng = np.random.default_rng(seed)
    op = VieOperator(dims, "PWL" if pwl else "PWC", 0.004)
    op.set_green_dense("N", bf.random_kernel(rng, dims, "N", pwl))
    nv = int(np.prod(dims))
    mat = vo.material(1 + 40 * rng.random(dims), rng.random(dims), vo.em_constants(128e6)["ce"])
    # a scatterer in a corner of the grid: the packed operators run on its support box
    inside = np.zeros(dims, bool)
    inside[: dims[0] // 2 + 1, : dims[1] // 2 + 1, :] = rng.random((dims[0] // 2 + 1, dims[1] // 2 + 1, dims[2])) < 0.8
    idxS = np.nonzero(inside.reshape(-1, order="F"))[0]
    op.set_material(mat["Mr"], mat["Mc"], idxS)
    return op, rng, nv


def _call(op, code, prec, x, out, nrhs):
    return op.lib.vie_apply_noalloc(op.h, code, prec, ctypes.c_void_p(x.data_ptr()),
                                    ctypes.c_void_p(out.data_ptr()), nrhs)


@pytest.mark.parametrize("opname,precision", [("VIE", "fp64"), ("N", "fp64"), ("VIE", "fp32"), ("VIEH", "fp32")])
def test_noalloc_matches_vie_apply_twice(opname, precision, monkeypatch):
    import torch
    from marie_vie import _lib as L
    codes = {"N": L.VIE_OP_N, "VIE": L.VIE_OP_VIE, "VIEH": L.VIE_OP_VIEH}
    prec = L.VIE_PREC_FP64 if precision == "fp64" else L.VIE_PREC_FP32
    cases = [((40, 36, 30), True, 2, 901), ((10, 9, 8), False, 1, 902)]   # the second input is smaller
    for dims, pwl, nrhs, seed in cases:
        print(f"case noalloc op={opname} {precision} dims={dims} pwl={pwl} nrhs={nrhs}", flush=True)
        op, rng, nv = _case(dims, pwl, seed)
        try:
            ql = 12 if pwl else 3
            n = ql * (op.nS if opname != "N" else nv)
            x = torch.from_numpy(rng.standard_normal((nrhs, n)) + 1j * rng.standard_normal((nrhs, n))).cuda()
            op.autotune(opname, precision, nrhs)          # set-up: tile widths, FP32 spectrum
            ref = op.apply(opname, x, precision)          # vie_apply, the context's own workspaces
            out = torch.full_like(x, complex(np.nan, np.nan))
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            monkeypatch.setenv("VIE_FAULT_INJECT_OOM", "1")
            try:
                rc = _call(op, codes[opname], prec, x, out, nrhs)
                torch.cuda.synchronize()
            finally:
                monkeypatch.delenv("VIE_FAULT_INJECT_OOM")
            assert rc == L.VIE_OK, op.lib.vie_last_error(op.h)
            assert torch.cuda.mem_get_info()[0] == free0
            assert torch.equal(out, ref)
        finally:
            op.close()


def test_noalloc_refuses_a_workspace_beyond_the_scratch():
    import torch
    from marie_vie import _lib as L
    dims = (10, 9, 8)
    op, rng, nv = _case(dims, True, 903)
    try:
        nrhs = 200              # N on the 20x18x16 grid: about 830 KB of workspace per rhs
        x = torch.zeros((nrhs, 12 * nv), dtype=torch.complex128, device="cuda")
        out = torch.zeros_like(x)
        rc = _call(op, L.VIE_OP_N, L.VIE_PREC_FP64, x, out, nrhs)
        torch.cuda.synchronize()
        assert rc == L.VIE_ERR_ARG
        assert b"exceeds the scratch" in op.lib.vie_last_error(op.h)
        # nothing was written
        assert not torch.any(out != 0)
        # FP32 before its spectrum copy exists: a state error, no allocation
        y = torch.zeros((1, 12 * nv), dtype=torch.complex128, device="cuda")
        rc = _call(op, L.VIE_OP_N, L.VIE_PREC_FP32, y, torch.zeros_like(y), 1)
        assert rc == L.VIE_ERR_STATE
    finally:
        op.close()
