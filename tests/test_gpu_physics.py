"""Physics sanity of the whole chain (Green tensor set-up, HIP operator,
device GMRES): the field inside a homogeneous dielectric sphere (radius
70 mm, 298 MHz, x-polarised plane wave) against the Mie series
(tests/mie.py, itself pinned by its Rayleigh limit 3/(eps+2) in
tests/test_mie.py).

The staircased sphere and the piecewise-constant basis limit the agreement
to discretisation level, so the checks are an accuracy bound at moderate
contrast (eps_r 15 on the 32^3 grid of 5 mm voxels: 0.086 measured) and, at
the head-like contrast of the C1 sphere (eps_r 50, sigma 0.6 S/m), first-
order convergence under refinement (0.368 at 5 mm, 0.146 at 2.5 mm measured
on MI355X).  A sign or scale error in the operator or the excitation leaves
an O(1) error that does not shrink with h.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FREQ = 298e6
RADIUS = 0.07


def _interior_error(dims, res, eps_r, sigma):
    import mie
    from marie_vie import VieOperator, green, phantoms
    ph = phantoms.sphere(dims=dims, res=res, radius=RADIUS, eps_r=eps_r, sigma=sigma)
    const = phantoms.constants(FREQ)
    op = VieOperator(ph.dims, "PWC", ph.res)
    try:
        op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
        Mr, Mc = ph.material(FREQ)
        op.set_material(Mr, Mc, ph.idxS)
        Ue = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
        mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
        b = const["ce"] * np.tile(mcr, 3) * Ue
        J, info = op.gmres(b, tol=1e-6, restart=50, maxit=400)
        assert info["resvec"][0][-1] <= 1e-6
    finally:
        op.close()
    nS = len(ph.idxS)
    E = (J.reshape(3, nS) / (const["ce"] * Mc.reshape(-1, order="F")[ph.idxS])).T
    cx, cy, cz = ph.centres()
    n1, n2 = ph.dims[0], ph.dims[1]
    s = ph.idxS
    pts = np.stack([cx[s % n1], cy[(s // n1) % n2], cz[s // (n1 * n2)]], axis=1)
    Em = mie.internal_field(pts, const["k0"], RADIUS, eps_r, sigma, const["omega"], phantoms.E0)
    inner = np.linalg.norm(pts, axis=1) < 0.8 * RADIUS
    err = np.linalg.norm(E[inner] - Em[inner]) / np.linalg.norm(Em[inner])
    print(f"sphere {dims} h={res * 1e3:.2f} mm eps_r={eps_r} sigma={sigma}: matvecs {info['nmatvec'][0]} "
          f"interior error vs Mie {err:.3f}", flush=True)
    return err


def test_sphere_moderate_contrast_matches_mie():
    assert _interior_error((32, 32, 32), 0.005, 15.0, 0.3) < 0.12


def test_c1_sphere_converges_to_mie():
    e_coarse = _interior_error((32, 32, 32), 0.005, 50.0, 0.6)
    e_fine = _interior_error((64, 64, 64), 0.0025, 50.0, 0.6)
    assert e_coarse < 0.45
    assert e_fine < 0.2
    assert e_fine < 0.5 * e_coarse
