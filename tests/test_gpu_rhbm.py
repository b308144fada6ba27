"""A measured body model through the whole chain: the reference's
inhomogeneous cylinder RHBM (downsampled 4x to 18 x 18 x 21 voxels of 8 mm,
tests/golden/cylinder_ds4.npz from tests/golden/make_rhbm_fixture.py), Green
tensor assembled on the GPU, plane-wave excitation at 298 MHz, device GMRES
against the oracle's GMRES on the same real-space kernel: the same [p outit],
the same resvec, the same solution; the mixed-precision solve reaches the
tolerance on the true residual."""
import os

import numpy as np
import pytest

from oracle import vie_oracle as vo

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def test_rhbm_cylinder_solve_matches_oracle():
    from marie_vie import VieOperator, green, phantoms
    d = np.load(os.path.join(GOLDEN, "cylinder_ds4.npz"))
    dims = tuple(int(v) for v in d["dims"])
    ph = phantoms.Phantom(dims, float(d["res"]), d["eps_r"], d["sigma"], d["idxS"],
                          tuple(float(v) for v in d["origin"]))
    freq = 298e6
    const = phantoms.constants(freq)
    op = VieOperator(dims, "PWC", ph.res)
    A = green.assemble(op, "N", const["k0"])
    op.set_green_dense("N", A)
    Mr, Mc = ph.material(freq)
    op.set_material(Mr, Mc, ph.idxS)
    Ue = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    b = const["ce"] * np.tile(mcr, 3) * Ue
    x, info = op.gmres(b, tol=1e-6, restart=50, maxit=200)
    xm, infm = op.gmres(b, tol=1e-6, restart=50, maxit=200, precision="mixed")
    rm = b - op.apply("VIE", xm)
    op.close()
    T = vo.circulant_dense(A[:, :, :, 0, :], False)       # PWC: (n1, n2, n3, 6)
    isol = vo.idx_solution(ph.idxS, int(np.prod(dims)), 3)
    afun = lambda v: vo.mvp_vie(v, Mr, Mc, isol, dims, ph.res, T, False)
    xr, itr, rvr = vo.gmres_vie(afun, b, restart=50, tol=1e-6, maxit=200)
    print(f"cylinder {dims} nS={len(ph.idxS)}: device iters {info['iters'][0]} oracle {list(itr)}; "
          f"mixed matvecs {infm['nmatvec'][0]}", flush=True)
    assert info["iters"][0] == list(itr)
    np.testing.assert_allclose(info["resvec"][0], rvr, rtol=1e-6, atol=1e-14)
    assert np.abs(x - xr).max() / np.abs(xr).max() < 1e-8
    assert np.linalg.norm(rm) <= 1e-6 * np.linalg.norm(b)
    assert np.abs(xm - x).max() / np.abs(x).max() < 1e-4


def test_config_builds_the_operator_of_a_body_file(tmp_path, monkeypatch):
    """A MARIE-style input (B0, Nucleus, basis, BodyFile) through
    marie_vie.config: the RHBM file written from the cylinder fixture, the
    operator built from it equals one set up directly, and its GMRES meets
    the configured tolerance."""
    print("case config -> operator", flush=True)
    monkeypatch.setenv("VIE_AXIS_FFT", "ct")      # both contexts on one FFT schedule
    from marie_vie import VieOperator, config, green, phantoms, rhbm
    d = np.load(os.path.join(GOLDEN, "cylinder_ds4.npz"))
    dims = tuple(int(v) for v in d["dims"])
    res = float(d["res"])
    o = np.array([float(v) for v in d["origin"]])
    g = np.stack(np.meshgrid(*[np.arange(n) for n in dims], indexing="ij"), axis=-1) * res + o
    m = rhbm.RHBM(g, d["eps_r"], d["sigma"], np.full(dims, 1000.0), np.asarray(d["idxS"], dtype=np.int64), "cyl")
    rhbm.save(str(tmp_path / "cyl.mat"), m)
    cfg = config.from_json({"B0": 7.0, "Nucleus": "1H", "Basis_Functions_VIE": 0, "BodyFile": "cyl.mat",
                            "tol": 1e-6})
    op, ph, const = config.build_operator(cfg, str(tmp_path))
    try:
        assert ph.dims == dims and abs(ph.res - res) < 1e-15
        rng = np.random.default_rng(8)
        x = rng.standard_normal(3 * len(ph.idxS)) + 1j * rng.standard_normal(3 * len(ph.idxS))
        y = op.apply("VIE", x)
        ref_op = VieOperator(dims, "PWC", res)
        ref_op.set_green_dense("N", green.assemble(ref_op, "N", const["k0"]))
        Mr, Mc = ph.material(cfg.freq_hz)
        ref_op.set_material(Mr, Mc, ph.idxS)
        np.testing.assert_array_equal(y, ref_op.apply("VIE", x))
        ref_op.close()
        Ue = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
        mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
        b = const["ce"] * np.tile(mcr, 3) * Ue
        xs, info = op.gmres(b, tol=cfg.tol, restart=cfg.restart, maxit=cfg.maxit, precision=cfg.precision)
        assert np.linalg.norm(b - op.apply("VIE", xs)) <= 1.01 * cfg.tol * np.linalg.norm(b)
    finally:
        op.close()
