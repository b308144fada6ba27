"""RHBM body-model I/O and downsampling (marie_vie.rhbm) against brute force.

The reference's two 2 mm models (data/bodies/Scatterers/2mm) are checked when
the reference tree is present; everything else uses synthetic models."""
import os

import numpy as np
import pytest

from marie_vie import rhbm

REF = "/root/reference/data/bodies/Scatterers/2mm"


def _synthetic(dims=(7, 5, 6), res=0.003, seed=3):
    rng = np.random.default_rng(seed)
    x0 = np.array([-0.01, 0.02, -0.005])
    lines = [x0[k] + np.arange(dims[k]) * res for k in range(3)]
    r = np.stack(np.meshgrid(*lines, indexing="ij"), axis=-1)
    eps = 1 + 70 * rng.random(dims)
    sig = rng.random(dims)
    rho = rng.integers(0, 200, dims).astype(np.uint8)
    idx = np.sort(rng.choice(int(np.prod(dims)), size=int(np.prod(dims)) // 3, replace=False))
    return rhbm.RHBM(r, eps, sig, rho, idx, "synthetic")


def test_save_load_round_trip(tmp_path):
    m = _synthetic()
    f = str(tmp_path / "m.mat")
    rhbm.save(f, m)
    g = rhbm.load(f)
    assert g.name == "synthetic" and g.dims == m.dims
    np.testing.assert_array_equal(g.idxS, m.idxS)
    for a, b in ((g.r, m.r), (g.epsilon_r, m.epsilon_r), (g.sigma_e, m.sigma_e), (g.rho, m.rho)):
        np.testing.assert_array_equal(a, b)
    assert abs(g.res - 0.003) < 1e-15
    ph = g.phantom()
    assert ph.dims == m.dims and ph.origin == tuple(m.r[0, 0, 0])
    np.testing.assert_allclose(ph.centres()[1], m.r[0, :, 0, 1])


@pytest.mark.parametrize("fov", ["crop", "pad"])
@pytest.mark.parametrize("policy", ["any", "all"])
def test_downsample_matches_brute_force(fov, policy):
    m = _synthetic()
    f = 2
    d = rhbm.downsample(m, f, mask_policy=policy, preserve_fov=fov)
    dims = m.dims
    if fov == "crop":
        n2 = [n // f for n in dims]
    else:
        n2 = [-(-n // f) for n in dims]
    assert d.dims == tuple(n2)
    mask = np.zeros(int(np.prod(dims)), bool)
    mask[m.idxS] = True
    mask = mask.reshape(dims, order="F")

    def fine(a, i, j, k):      # edge-replicated element (the 'pad' rule)
        return a[min(i, dims[0] - 1), min(j, dims[1] - 1), min(k, dims[2] - 1)]

    m2 = np.zeros(n2, bool)
    for I in range(n2[0]):
        for J in range(n2[1]):
            for K in range(n2[2]):
                cells = [(f * I + a, f * J + b, f * K + c) for a in range(f) for b in range(f) for c in range(f)]
                ev = np.mean([fine(m.epsilon_r, *t) for t in cells])
                assert abs(d.epsilon_r[I, J, K] - ev) < 1e-12
                rv = np.mean([float(fine(m.rho, *t)) for t in cells])
                assert d.rho[I, J, K] == np.floor(rv + 0.5)        # uint8 cast rounds
                mm = [fine(mask, *t) for t in cells]
                m2[I, J, K] = any(mm) if policy == "any" else all(mm)
    np.testing.assert_array_equal(d.idxS, np.nonzero(m2.reshape(-1, order="F"))[0])
    np.testing.assert_allclose(d.r[1, 0, 0, 0] - d.r[0, 0, 0, 0], f * 0.003, rtol=1e-12)
    np.testing.assert_allclose(d.r[0, 0, 0], m.r[0, 0, 0])


def test_nearest_takes_every_factor_th_voxel():
    m = _synthetic()
    d = rhbm.downsample(m, 2, use_nearest=True)
    np.testing.assert_array_equal(d.epsilon_r, m.epsilon_r[:6:2, :4:2, :6:2])


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data not present")
def test_reference_models_load_and_downsample():
    cyl = rhbm.load(os.path.join(REF, "inhomogeneous_cylinder.mat"))
    assert cyl.dims == (73, 73, 84) and abs(cyl.res - 0.002) < 1e-12 and len(cyl.idxS) == 316470
    sph = rhbm.load(os.path.join(REF, "Multilayer_Sphere.mat"))
    assert sph.dims == (121, 121, 121) and len(sph.idxS) == 904089
    d = rhbm.downsample(cyl, 2)
    assert d.dims == (36, 36, 42) and abs(d.res - 0.004) < 1e-12
    mask = np.zeros(73 * 73 * 84, bool)
    mask[cyl.idxS] = True
    blk = mask.reshape((73, 73, 84), order="F")[:72, :72, :84].reshape((2, 36, 2, 36, 2, 42), order="F")
    assert len(d.idxS) == int(blk.any(axis=(0, 2, 4)).sum())
