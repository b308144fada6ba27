"""Physics sanity of the whole chain (Green tensor set-up, HIP operator,
device GMRES): the field inside the C1 dielectric sphere (32^3 grid of 5 mm
voxels, radius 70 mm, eps_r 50, sigma 0.6 S/m, 298 MHz, x-polarised plane
wave) against the Mie series (tests/mie.py).  The staircased sphere limits
the agreement to discretisation level, so the bound is loose; the Mie check
itself is pinned by its Rayleigh limit 3/(eps+2) (tests/test_mie.py)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_c1_sphere_field_matches_mie():
    import mie
    from marie_vie import VieOperator, green, phantoms
    freq = 298e6
    ph = phantoms.sphere()
    const = phantoms.constants(freq)
    op = VieOperator(ph.dims, "PWC", ph.res)
    op.set_green_dense("N", green.assemble(op, "N", const["k0"]))
    Mr, Mc = ph.material(freq)
    op.set_material(Mr, Mc, ph.idxS)
    Ue = phantoms.project(ph, phantoms.plane_wave(const["k0"]), False)
    mcr = (Mc.reshape(-1, order="F") / Mr.reshape(-1, order="F"))[ph.idxS]
    b = const["ce"] * np.tile(mcr, 3) * Ue
    J, info = op.gmres(b, tol=1e-6, restart=50, maxit=200)
    print(f"C1 sphere: iters {info['iters'][0]} matvecs {info['nmatvec'][0]}", flush=True)
    assert info["resvec"][0][-1] <= 1e-6
    nS = len(ph.idxS)
    E = (J.reshape(3, nS) / (const["ce"] * Mc.reshape(-1, order="F")[ph.idxS])).T
    cx, cy, cz = ph.centres()
    n1, n2 = ph.dims[0], ph.dims[1]
    s = ph.idxS
    pts = np.stack([cx[s % n1], cy[(s // n1) % n2], cz[s // (n1 * n2)]], axis=1)
    Em = mie.internal_field(pts, const["k0"], 0.07, 50.0, 0.6, const["omega"], phantoms.E0)
    inner = np.linalg.norm(pts, axis=1) < 0.8 * 0.07
    err = np.linalg.norm(E[inner] - Em[inner]) / np.linalg.norm(Em[inner])
    print(f"C1 sphere: interior field relative error vs Mie {err:.3f}", flush=True)
    assert err < 0.15
    op.close()
