"""Run configuration (marie_vie.config): the Larmor frequency of em_constants.m
and the input JSON and defaults of load_inputs.m."""
import json
import math

import pytest

from marie_vie import config

REF_INPUT = {"B0": 3.0063, "Nucleus": "1H", "Basis_Functions_VIE": 1,
             "BodyFile": "Scatterers/2mm/Multilayer_Sphere.mat", "CoilFile": "",
             "WireFile": "Head_Receivers/array.msh", "ShieldFile": "", "SurfaceBasisSupportFile": "",
             "TMD": 1, "BasisFile": "basis.mat"}


def test_larmor_frequency():
    # em_constants.m:41-71: freq = |gamma0| B0
    f = config.larmor_hz(3.0063, "1H")
    assert f == 42.577478518e6 * 3.0063
    # the 3 T head workload: 128.0 MHz, k0 = 2 pi f / c0 = 2.6827 rad/m
    assert f == pytest.approx(128.0e6, rel=1e-3)
    assert 2 * math.pi * f / 299792458.0 == pytest.approx(2.6827, abs=5e-5)
    assert config.larmor_hz(7.0) == pytest.approx(298.04e6, rel=1e-5)
    assert config.larmor_hz(3.0, "13c") == pytest.approx(32.115e6, rel=1e-9)
    assert config.larmor_hz(1.5, "129Xe") == pytest.approx(17.6655e6, rel=1e-9)
    with pytest.raises(ValueError):
        config.larmor_hz(3.0, "7Li")


def test_reference_input_json(tmp_path):
    p = tmp_path / "inp.json"
    p.write_text(json.dumps(REF_INPUT))
    for src in (REF_INPUT, json.dumps(REF_INPUT), str(p)):
        c = config.from_json(src)
        assert c.basis == "PWL" and c.freq_hz == 42.577478518e6 * 3.0063
        # load_inputs.m:68-81 numerics
        assert c.tol == 1e-5 and c.tucker_tol == pytest.approx(1e-7) and c.far_nq == 4 and c.medium_nq == 8
        assert c.restart == 50 and c.maxit == 200 and c.precision == "fp64"
        assert c.body_file == REF_INPUT["BodyFile"] and c.other_files["WireFile"] == REF_INPUT["WireFile"]


def test_overrides_and_errors():
    c = config.from_json(dict(REF_INPUT, Basis_Functions_VIE=0, freq_hz=298e6, tol=1e-6, precision="mixed"))
    assert c.basis == "PWC" and c.freq_hz == 298e6 and c.tol == 1e-6 and c.precision == "mixed"
    assert c.tucker_tol == pytest.approx(1e-8)
    bad = [dict(REF_INPUT, Basis_Functions_VIE=2), {k: v for k, v in REF_INPUT.items() if k != "B0"},
           dict(REF_INPUT, precision="fp16"), dict(REF_INPUT, Nucleus="7Li"), dict(REF_INPUT, tol=0)]
    for d in bad:
        with pytest.raises(ValueError):
            config.from_json(d)
