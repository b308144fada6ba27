"""Host-side handle of the HIP VIE operator (one context per voxel grid).

``VieOperator`` owns a libvie_hip context: the circulant Green spectra (N
and/or K, FP64 on the device), the material contrast and the workspaces.  It
accepts numpy arrays (host memory; the call is synchronous) or torch
complex128 CUDA tensors (device memory; the call is queued on torch's
current stream: every torch-tensor call first points the context at the
stream torch.cuda.current_stream() returns at that moment, so the kernels
run after the work that produced the inputs and before any later work on
that stream reads the outputs).

Layouts follow MATLAB (the reference's) conventions:
  dense field  : [n1, n2, n3, ql] column-major, i.e. numpy shape (n1,n2,n3,ql)
                 in Fortran order, or a C-contiguous torch tensor of shape
                 (nrhs, ql, n3, n2, n1)
  packed vector: ql*nS values ordered like idx_solution
                 (geo_body_domain.m:20-35); numpy (ql*nS,) or (ql*nS, nrhs)
                 Fortran order, torch (nrhs, ql*nS) C-contiguous
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib as L

_OPS = {"N": L.VIE_OP_N, "K": L.VIE_OP_K, "NH": L.VIE_OP_NH, "KH": L.VIE_OP_KH,
        "VIE": L.VIE_OP_VIE, "VIEH": L.VIE_OP_VIEH}
_PREC = {"fp64": L.VIE_PREC_FP64, "fp32": L.VIE_PREC_FP32, "mixed": L.VIE_PREC_FP32}


def _ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


class VieOperator:
    """Device-resident VIE operator for one grid (n1, n2, n3), PWC or PWL basis."""

    def __init__(self, dims, basis: str, res: float, device: int = 0):
        self.lib = L.load()
        self.dims = tuple(int(d) for d in dims)
        self.pwl = basis.upper() == "PWL"
        self.ql = 12 if self.pwl else 3
        self.res = float(res)
        desc = L.VieDesc(self.dims[0], self.dims[1], self.dims[2],
                         L.VIE_BASIS_PWL if self.pwl else L.VIE_BASIS_PWC, self.res, device)
        h = ctypes.c_void_p()
        rc = self.lib.vie_create(ctypes.byref(desc), ctypes.byref(h))
        L.check(rc, None)
        self.h = h
        self.nS = 0
        self._stream = None   # HIP stream the context runs on, when set from torch
        P = [ctypes.c_int32() for _ in range(3)]
        self.lib.vie_fft_dims(self.h, *[ctypes.byref(p) for p in P])
        self.fft_dims = tuple(p.value for p in P)

    # -- life cycle -------------------------------------------------------
    def close(self):
        if getattr(self, "h", None):
            self.lib.vie_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, rc):
        L.check(rc, self.h)

    def use_torch_stream(self):
        """Queue device-pointer calls on torch's current HIP stream (torch-tensor
        calls do this by themselves)."""
        import torch
        s = torch.cuda.current_stream().cuda_stream
        if s != self._stream:
            self._check(self.lib.vie_set_stream(self.h, ctypes.c_void_p(s)))
            self._stream = s

    def set_gram(self, res: float):
        """Voxel edge of the Gram diagonal (invG) the VIE operators apply
        (mvp_invG_pwl.m:3-7); N and K do not depend on it."""
        self._check(self.lib.vie_set_gram(self.h, float(res)))
        self.res = float(res)

    def autotune(self, op: str = "VIE", precision: str = "fp64", nrhs: int = 1):
        """Tile widths of the z-forward, y-forward, y-inverse and z-inverse
        axis passes for (op, precision, nrhs), tuned now if the context has
        not fixed them yet for this precision and batch shape."""
        tiles = np.zeros(4, dtype=np.int32)
        self._check(self.lib.vie_autotune(self.h, _OPS[op], _PREC[precision], int(nrhs), _ptr(tiles)))
        return tiles.tolist()

    def schedule(self, precision: str = "fp64"):
        """FFT schedule of the z-forward, y-forward, y-inverse and z-inverse
        axis passes: "ct" (Stockham stages, largest radix first), "rf"
        (two-step register split), "cto" (Stockham stages, odd radices first)
        or None where not fixed yet."""
        sched = np.zeros(4, dtype=np.int32)
        self._check(self.lib.vie_axis_schedule(self.h, _PREC[precision], _ptr(sched)))
        return [{0: None, 1: "ct", 2: "rf", 3: "cto"}[int(v)] for v in sched]

    def box(self):
        """Support box of the packed operators (VIE, VIEH, GMRES): origin,
        extent and FFT lengths of the scatterer's bounding box when the
        operators run on it, None when they run on the full grid."""
        lo, w, P = (np.zeros(3, dtype=np.int32) for _ in range(3))
        self._check(self.lib.vie_box_dims(self.h, _ptr(lo), _ptr(w), _ptr(P)))
        if not w.any():
            return None
        return {"lo": tuple(int(v) for v in lo), "extent": tuple(int(v) for v in w),
                "fft": tuple(int(v) for v in P)}

    # -- set-up ---------------------------------------------------------------
    def set_green_dense(self, op: str, A: np.ndarray):
        """Real-space kernel A[n1,n2,n3,pwx,ncomp] (assembly_N.m:16 layout)."""
        A = np.asarray(A, dtype=np.complex128)
        if A.ndim == 4:                       # PWC: (n1,n2,n3,ncomp)
            A = A[:, :, :, None, :]
        pwx = 10 if self.pwl else 1
        ncomp = 6 if op == "N" else 3
        if A.shape != self.dims + (pwx, ncomp):
            raise ValueError(f"A has shape {A.shape}, expected {self.dims + (pwx, ncomp)}")
        Af = np.asfortranarray(A)
        self._check(self.lib.vie_set_green_dense(self.h, _OPS[op], _ptr(Af), L.VIE_MEM_HOST))

    def set_green_tucker(self, op: str, T):
        """Reference T-struct (assembly_fft_circ_tucker_*.m): PWC list T[j],
        PWL nested list T[i][j]; each term has G, U1, U2, U3 (numpy)."""
        ncomp = 6 if op == "N" else 3
        terms = []
        if self.pwl:
            for i in range(10):
                for j in range(ncomp):
                    terms.append((i * ncomp + j, T[i][j]))
        else:
            for j in range(ncomp):
                terms.append((j, T[j]))
        for g, t in terms:
            G = np.asfortranarray(np.asarray(t.G, dtype=np.complex128))
            if G.ndim < 3:
                G = G.reshape(G.shape + (1,) * (3 - G.ndim), order="F")
            U = [np.asfortranarray(np.asarray(u, dtype=np.complex128)) for u in (t.U1, t.U2, t.U3)]
            r1, r2, r3 = G.shape
            self._check(self.lib.vie_set_green_tucker(self.h, _OPS[op], g, _ptr(G), r1, r2, r3,
                                                      _ptr(U[0]), _ptr(U[1]), _ptr(U[2])))

    def set_material(self, Mr: np.ndarray, Mc: np.ndarray, idxS: np.ndarray):
        """Mr, Mc on the (n1,n2,n3) grid; idxS 0-based Fortran-order scatterer voxels."""
        Mr = np.asfortranarray(np.asarray(Mr, dtype=np.complex128).reshape(self.dims, order="F"))
        Mc = np.asfortranarray(np.asarray(Mc, dtype=np.complex128).reshape(self.dims, order="F"))
        idx = np.ascontiguousarray(np.asarray(idxS, dtype=np.int64))
        self._check(self.lib.vie_set_material(self.h, _ptr(Mr), _ptr(Mc), _ptr(idx), idx.size))
        self.nS = int(idx.size)

    # -- application ------------------------------------------------------------
    def packed_len(self) -> int:
        return self.ql * self.nS

    def apply(self, op: str, x, precision: str = "fp64", out=None):
        """Apply op in {N, K, NH, KH, VIE, VIEH} to x (numpy or torch CUDA).
        For torch input an output tensor of the same shape may be passed."""
        code = _OPS[op]
        prec = _PREC[precision]
        packed = op in ("VIE", "VIEH")
        per = self.packed_len() if packed else self.ql * int(np.prod(self.dims))
        if _is_torch(x):
            import torch
            if not x.is_cuda or x.dtype != torch.complex128 or not x.is_contiguous():
                raise ValueError("torch input must be a contiguous complex128 CUDA tensor")
            k = x.numel() // per
            if k * per != x.numel():
                raise ValueError("input size is not a multiple of the vector length")
            if out is None:
                out = torch.empty_like(x)
            elif out.shape != x.shape or out.dtype != x.dtype or not out.is_contiguous():
                raise ValueError("out must match x")
            self.use_torch_stream()
            self._check(self.lib.vie_apply(self.h, code, prec, ctypes.c_void_p(x.data_ptr()),
                                           ctypes.c_void_p(out.data_ptr()), k, L.VIE_MEM_DEVICE))
            return out
        xa = np.asarray(x, dtype=np.complex128)
        shape = xa.shape
        xf = np.asfortranarray(xa)
        k = xf.size // per
        if k * per != xf.size:
            raise ValueError(f"input of size {xf.size} is not a multiple of {per}")
        out = np.empty_like(xf, order="F")
        self._check(self.lib.vie_apply(self.h, code, prec, _ptr(xf), _ptr(out), k, L.VIE_MEM_HOST))
        return out.reshape(shape, order="F")

    # -- solver ---------------------------------------------------------------------
    def gmres(self, b, x0=None, tol=1e-6, restart=50, maxit=200, precision="fp64", adjoint=False):
        """Device GMRES on the VIE operator (is_gmres_vie.m semantics for fp64).

        b: packed (ql*nS,) or (ql*nS, nrhs) numpy, or torch (nrhs, ql*nS) CUDA.
        Returns (x, info) with info = {iters: [[p, outit]...], resvec: [...],
        nmatvec: [...]}."""
        op = L.VIE_OP_VIEH if adjoint else L.VIE_OP_VIE
        prec = _PREC[precision]
        n = self.packed_len()
        if n == 0:
            raise ValueError("set_material first (no scatterer voxels)")
        cap = restart * maxit + 1
        if _is_torch(b):
            import torch
            if not b.is_cuda or b.dtype != torch.complex128 or not b.is_contiguous():
                raise ValueError("torch b must be a contiguous complex128 CUDA tensor")
            nrhs = b.numel() // n
            if nrhs * n != b.numel():
                raise ValueError(f"b of size {b.numel()} is not a multiple of {n}")
            if x0 is not None and (x0.shape != b.shape or x0.dtype != b.dtype or not x0.is_cuda):
                raise ValueError("x0 must match b")
            x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
            self.use_torch_stream()
            bptr, xptr, mem = b.data_ptr(), x.data_ptr(), L.VIE_MEM_DEVICE
            keep = (b, x)
        else:
            ba = np.asfortranarray(np.asarray(b, dtype=np.complex128))
            nrhs = ba.size // n
            if nrhs * n != ba.size:
                raise ValueError(f"b of size {ba.size} is not a multiple of {n}")
            if x0 is None:
                x = np.zeros_like(ba, order="F")
            else:
                x = np.asfortranarray(np.array(x0, dtype=np.complex128))
                if x.shape != ba.shape:
                    raise ValueError("x0 must match b")
            bptr, xptr, mem = ba.ctypes.data, x.ctypes.data, L.VIE_MEM_HOST
            keep = (ba, x)
        resvec = np.zeros((nrhs, cap))
        nres = np.zeros(nrhs, dtype=np.int32)
        iters = np.zeros((nrhs, 2), dtype=np.int32)
        nmv = np.zeros(nrhs, dtype=np.int32)
        rc = self.lib.vie_gmres(self.h, op, prec, ctypes.c_void_p(bptr), ctypes.c_void_p(xptr), nrhs,
                                float(tol), int(restart), int(maxit), mem, _ptr(resvec), _ptr(nres),
                                _ptr(iters), _ptr(nmv))
        del keep
        self._check(rc)
        info = {"iters": iters.tolist(), "nmatvec": nmv.tolist(),
                "resvec": [resvec[i, :nres[i]].copy() for i in range(nrhs)]}
        return x, info

    def cocr(self, b, x0=None, tol=1e-6, maxit=1000, precision="fp64", adjoint=False):
        """Device COCR on the VIE operator (vie_cocr): the complex-symmetric
        short-recurrence solver; precision "mixed" (or "fp32") runs the FP32
        operator with FP64 recurrence scalars and FP64 reliable updates.

        b: as for gmres.  Returns (x, info) with info = {nmatvec: [...],
        nreplace: [...] (FP64 residual replacements), resvec: [...]} per rhs;
        the last resvec entry of a converged rhs is its FP64 residual."""
        return self._sym_solve(self.lib.vie_cocr, b, x0, tol, maxit, precision, adjoint)

    def block_cocg(self, b, x0=None, tol=1e-6, maxit=1000, precision="fp64", adjoint=False):
        """Device block COCG on the VIE operator (vie_bcocg): the rhs share
        one block Krylov space (groups of up to 32), so a batch needs fewer
        operator applications per rhs than COCR or GMRES one rhs at a time.
        Arguments and info as cocr; nmatvec counts per rhs the block
        iterations plus FP64 residual evaluations."""
        return self._sym_solve(self.lib.vie_bcocg, b, x0, tol, maxit, precision, adjoint)

    def _sym_solve(self, fn, b, x0, tol, maxit, precision, adjoint):
        op = L.VIE_OP_VIEH if adjoint else L.VIE_OP_VIE
        prec = _PREC[precision]
        n = self.packed_len()
        if n == 0:
            raise ValueError("set_material first (no scatterer voxels)")
        cap = 2 * maxit + 2
        if _is_torch(b):
            import torch
            if not b.is_cuda or b.dtype != torch.complex128 or not b.is_contiguous():
                raise ValueError("torch b must be a contiguous complex128 CUDA tensor")
            nrhs = b.numel() // n
            if nrhs * n != b.numel():
                raise ValueError(f"b of size {b.numel()} is not a multiple of {n}")
            if x0 is not None and (x0.shape != b.shape or x0.dtype != b.dtype or not x0.is_cuda):
                raise ValueError("x0 must match b")
            x = torch.zeros_like(b) if x0 is None else x0.clone().contiguous()
            self.use_torch_stream()
            bptr, xptr, mem = b.data_ptr(), x.data_ptr(), L.VIE_MEM_DEVICE
            keep = (b, x)
        else:
            ba = np.asfortranarray(np.asarray(b, dtype=np.complex128))
            nrhs = ba.size // n
            if nrhs * n != ba.size:
                raise ValueError(f"b of size {ba.size} is not a multiple of {n}")
            if x0 is None:
                x = np.zeros_like(ba, order="F")
            else:
                x = np.asfortranarray(np.array(x0, dtype=np.complex128))
                if x.shape != ba.shape:
                    raise ValueError("x0 must match b")
            bptr, xptr, mem = ba.ctypes.data, x.ctypes.data, L.VIE_MEM_HOST
            keep = (ba, x)
        resvec = np.zeros((nrhs, cap))
        nres = np.zeros(nrhs, dtype=np.int32)
        nmv = np.zeros(nrhs, dtype=np.int32)
        nrep = np.zeros(nrhs, dtype=np.int32)
        rc = fn(self.h, op, prec, ctypes.c_void_p(bptr), ctypes.c_void_p(xptr), nrhs, float(tol), int(maxit), mem,
                _ptr(resvec), _ptr(nres), _ptr(nmv), _ptr(nrep))
        del keep
        self._check(rc)
        info = {"nmatvec": nmv.tolist(), "nreplace": nrep.tolist(),
                "resvec": [resvec[i, :nres[i]].copy() for i in range(nrhs)]}
        return x, info
