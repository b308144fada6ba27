"""Writes tests/golden/cylinder_ds4.npz: the reference's experimental
inhomogeneous cylinder body model (data/bodies/Scatterers/2mm, MAT v5 data)
downsampled 4x (block mean, mask 'any', crop) to 18 x 18 x 21 voxels of
8 mm, for the GPU solve test on a measured body model.  Run here, where the
reference tree exists:  python tests/golden/make_rhbm_fixture.py"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(HERE, "..", "..", "marie-tools_amd"))
from marie_vie import rhbm  # noqa: E402

SRC = "/root/reference/data/bodies/Scatterers/2mm/inhomogeneous_cylinder.mat"

m = rhbm.downsample(rhbm.load(SRC), 4)
ph = m.phantom()
np.savez_compressed(os.path.join(HERE, "cylinder_ds4.npz"), eps_r=ph.eps_r, sigma=ph.sigma, idxS=ph.idxS,
                    origin=np.array(ph.origin), res=np.array(ph.res), dims=np.array(ph.dims))
print(ph.dims, ph.res, len(ph.idxS))
