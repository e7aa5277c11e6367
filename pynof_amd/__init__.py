"""pynof_amd: the density-fitted JK build of PyNOF on AMD MI355X (gfx950).

Entry points:
  computeJKj, computeJKH_MO   drop-in replacements of PyNOF's RI JK builds
  JKH_MO_tmp                  drop-in for the orbital-gradient MO integrals (RI)
  computeLagrange2            drop-in for the orbital-gradient Lagrangian (RI, closed shell)
  computeF_RC_driver, computeF_RO_driver, computeLagrange, ENERGY1r,
  computeLagrangeConvergency  drop-ins for the Fock route and the orbital step's energy call
  ocupacion, PNOFi_selector, calce, calcocce, calcoccg
                              drop-ins for the occupation step (pnof.py)
  JKEngine                    device-resident b and fused builds
"""
from ._lib import PnjkError  # noqa: F401
from .engine import JKEngine  # noqa: F401
from .integrals import (  # noqa: F401
    ENERGY1r,
    JKH_MO_tmp,
    computeLagrangeConvergency,
    computeF_RC_driver,
    computeF_RO_driver,
    computeJKH_MO,
    computeJKj,
    computeLagrange,
    computeLagrange2,
)
from .parameters import param  # noqa: F401
from .pnof import PNOFi_selector, calce, calcocce, calcoccg, ocupacion  # noqa: F401

__all__ = ["JKEngine", "computeJKj", "computeJKH_MO", "JKH_MO_tmp", "computeLagrange2", "computeF_RC_driver",
           "computeF_RO_driver", "computeLagrange", "ENERGY1r", "computeLagrangeConvergency", "ocupacion",
           "PNOFi_selector", "calce", "calcocce", "calcoccg", "param", "PnjkError"]
