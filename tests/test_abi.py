"""C-ABI checks that need no GPU: the library loads, exports every entry
point include/vie_op.h declares, validates arguments before touching the
device, and reports allocation failure as VIE_ERR_OOM (the status a MATLAB
gateway turns into an error so the caller's GPU -> CPU fallback in
solver_wsvie.m:10-23 keeps working)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    src = open(os.path.join(ROOT, "include", "vie_op.h")).read()
    return sorted(set(re.findall(r"\b(vie_[a-z_0-9]+)\s*\(", src)))


def test_header_symbols_exported():
    from marie_vie import _lib as L
    lib = L.load()
    names = _declared()
    assert len(names) >= 12
    for n in names:
        assert hasattr(lib, n), f"{n} declared in vie_op.h but not exported"
    assert set(names) == set(L.EXPORTS)


def test_one_hip_runtime_per_process():
    """The package loads the library after torch, so the process maps one HIP
    runtime (torch's) whichever of torch and the library a caller touches
    first; two runtimes leave torch without a device."""
    import ast
    import subprocess
    import sys
    code = ("import sys; sys.path[:0] = [%r, %r]\n"
            "from marie_vie import _lib\n_lib.load()\nimport torch\n"
            "maps = open('/proc/self/maps').read().splitlines()\n"
            "print(sorted({l.split()[-1] for l in maps if 'libamdhip64' in l}))\n"
            % (ROOT, os.path.join(ROOT, "marie-tools_amd")))
    out = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    paths = ast.literal_eval(out.stdout.strip().splitlines()[-1])
    assert len(paths) == 1, paths


def test_invalid_descriptor_rejected_without_device():
    from marie_vie import _lib as L
    lib = L.load()
    h = ctypes.c_void_p()
    desc = L.VieDesc(0, 4, 4, L.VIE_BASIS_PWC, 0.005, 0)
    assert lib.vie_create(ctypes.byref(desc), ctypes.byref(h)) == L.VIE_ERR_ARG
    assert b"dimensions" in lib.vie_last_error(None)
    desc = L.VieDesc(4, 4, 4, 7, 0.005, 0)
    assert lib.vie_create(ctypes.byref(desc), ctypes.byref(h)) == L.VIE_ERR_ARG
    desc = L.VieDesc(4, 4, 4, L.VIE_BASIS_PWL, -1.0, 0)
    assert lib.vie_create(ctypes.byref(desc), ctypes.byref(h)) == L.VIE_ERR_ARG
    assert lib.vie_create(None, ctypes.byref(h)) == L.VIE_ERR_ARG


def test_null_handle_calls_return_arg_error():
    from marie_vie import _lib as L
    lib = L.load()
    assert lib.vie_apply(None, 0, 0, None, None, 1, 0) == L.VIE_ERR_ARG
    assert lib.vie_gmres(None, 4, 0, None, None, 1, 1e-6, 50, 200, 0, None, None, None, None) == L.VIE_ERR_ARG
    assert lib.vie_set_material(None, None, None, None, 0) == L.VIE_ERR_ARG
    assert lib.vie_destroy(None) == L.VIE_OK
    assert lib.vie_num_scatterers(None) == -1


def test_allocation_failure_is_reported_as_oom(monkeypatch):
    from marie_vie import _lib as L
    from marie_vie import VieOperator, VieError
    monkeypatch.setenv("VIE_FAULT_INJECT_OOM", "1")
    with pytest.raises(VieError) as ei:
        VieOperator((8, 8, 8), "PWL", 0.004)
    assert ei.value.code == L.VIE_ERR_OOM
    assert "VIE_FAULT_INJECT_OOM" in str(ei.value)


def test_unsupported_lds_size_rejected():
    from marie_vie import _lib as L
    lib = L.load()
    h = ctypes.c_void_p()
    # PWL with n1 = 300 needs 2*12*600 complex values of LDS in the fused x pass
    desc = L.VieDesc(300, 4, 4, L.VIE_BASIS_PWL, 0.004, 0)
    assert lib.vie_create(ctypes.byref(desc), ctypes.byref(h)) == L.VIE_ERR_ARG
