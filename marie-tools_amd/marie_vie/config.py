"""Run configuration of the VIE solve: the part of MARIE's input that reaches
the operator and the GMRES.

MARIE reads a JSON input (load_inputs.m:48-62: B0, Nucleus,
Basis_Functions_VIE 0 = PWC / 1 = PWL, BodyFile, coil/wire/shield files) and
hard-codes the numerics (load_inputs.m:68-99: GMRES tol 1e-5, HOSVD tol
tol * 1e-2, Gauss orders 4 / 8 for the far / medium offsets); the frequency is the Larmor frequency
|gamma0| B0 of the nucleus (em_constants.m:41-78).  VieConfig holds the same
fields with those defaults, lets a run state the frequency directly and
override tol, restart, maxit, precision and the Tucker tolerance (the
benchmark workloads solve to 1e-6), and build_operator() sets up the device
operator for the configured RHBM body.  Coil, wire and shield files are kept
as read but not used: the coupled solvers they drive are not part of this
library.
"""
from __future__ import annotations

import json
import os
from dataclasses import dataclass, field

#: gyromagnetic ratio / (2 pi) in Hz/T per nucleus (em_constants.m:41-64)
GAMMA0_HZ_PER_T = {"1H": 42.577478518e6, "2H": 6.536e6, "3HE": 32.434e6, "13C": 10.705e6, "129XE": 11.777e6}


def larmor_hz(b0_tesla: float, nucleus: str = "1H") -> float:
    """|gamma0| B0 (em_constants.m:71)."""
    key = nucleus.upper()
    if key not in GAMMA0_HZ_PER_T:
        raise ValueError(f"unknown nucleus {nucleus!r}; known: {sorted(GAMMA0_HZ_PER_T)}")
    return abs(GAMMA0_HZ_PER_T[key]) * float(b0_tesla)


@dataclass
class VieConfig:
    freq_hz: float
    basis: str = "PWL"                 # "PWC" or "PWL"
    body_file: str = ""
    tol: float = 1e-5                  # load_inputs.m:68
    tucker_tol: float = 1e-7           # tol * 1e-2, load_inputs.m:69
    restart: int = 50                  # is_gmres_vie callers (ie_solver_vie.m:19)
    maxit: int = 200
    precision: str = "fp64"            # or "mixed"
    far_nq: int = 4                    # load_inputs.m:79-80 (the near field uses the library's own rule)
    medium_nq: int = 8
    b0_tesla: float | None = None
    nucleus: str = "1H"
    other_files: dict = field(default_factory=dict)

    def __post_init__(self):
        if self.basis not in ("PWC", "PWL"):
            raise ValueError("basis must be 'PWC' or 'PWL'")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if not (self.freq_hz > 0 and self.tol > 0 and self.restart >= 1 and self.maxit >= 1):
            raise ValueError("freq_hz, tol, restart and maxit must be positive")


def from_json(src) -> VieConfig:
    """A MARIE input JSON (path, text or dict), with optional overrides
    freq_hz, tol, tucker_tol, restart, maxit, precision.  The frequency is
    freq_hz when given, else the Larmor frequency of B0 and Nucleus."""
    if isinstance(src, dict):
        d = dict(src)
    elif isinstance(src, str) and src.lstrip().startswith("{"):
        d = json.loads(src)
    else:
        with open(src) as f:
            d = json.load(f)
    nucleus = d.get("Nucleus", "1H")
    b0 = d.get("B0")
    if "freq_hz" in d:
        freq = float(d["freq_hz"])
    elif b0 is not None:
        freq = larmor_hz(float(b0), nucleus)
    else:
        raise ValueError("the input needs B0 (with Nucleus) or freq_hz")
    pwx = int(d.get("Basis_Functions_VIE", 1))
    if pwx not in (0, 1):
        raise ValueError("Basis_Functions_VIE must be 0 (PWC) or 1 (PWL)")
    tol = float(d.get("tol", 1e-5))
    others = {k: d[k] for k in ("CoilFile", "WireFile", "ShieldFile", "SurfaceBasisSupportFile", "BasisFile", "TMD")
              if k in d}
    return VieConfig(freq_hz=freq, basis="PWL" if pwx == 1 else "PWC", body_file=d.get("BodyFile", ""),
                     tol=tol, tucker_tol=float(d.get("tucker_tol", tol * 1e-2)),
                     restart=int(d.get("restart", 50)), maxit=int(d.get("maxit", 200)),
                     precision=d.get("precision", "fp64"), b0_tesla=None if b0 is None else float(b0),
                     nucleus=nucleus, other_files=others)


def build_operator(cfg: VieConfig, bodies_dir: str = "."):
    """The device operator of the configured body: RHBM read (geo_body_domain.m),
    Green tensor assembled at the configured orders, material set.  Returns
    (op, phantom, constants); the caller closes op."""
    from . import green, phantoms, rhbm
    from .operator import VieOperator
    path = cfg.body_file if os.path.isabs(cfg.body_file) else os.path.join(bodies_dir, cfg.body_file)
    ph = rhbm.load(path).phantom()
    const = phantoms.constants(cfg.freq_hz)
    op = VieOperator(ph.dims, cfg.basis, ph.res)
    try:
        op.set_green_dense("N", green.assemble(op, "N", const["k0"], far_nq=cfg.far_nq, medium_nq=cfg.medium_nq))
        Mr, Mc = ph.material(cfg.freq_hz)
        op.set_material(Mr, Mc, ph.idxS)
    except Exception:
        op.close()
        raise
    return op, ph, const
