"""Drop-in replacements for the occupation step of PyNOF (src/pynof/pnof.py).

``ocupacion``, ``PNOFi_selector``, ``calce``, ``calcocce`` and ``calcoccg`` keep the
reference's signatures, argument meaning and return values (pnof.py:692, :657,
:905, :994, :1004) and run on the MI355X through the HIP library; there is no CPU
path. Supported: PNOF5, PNOF7 and GNOF (p.ipnof 5, 7, 8), the Softmax and
Trigonometric parametrisations, closed (MSpin = 0) and open shell. Other
settings raise NotImplementedError. Arrays may be numpy (host results) or torch device tensors
(device results).
"""
from __future__ import annotations

import numpy as np

from .engine import JKEngine
from .integrals import _plain_engine


def _pnof_of(p, occ_method=None):
    ipnof = int(getattr(p, "ipnof", 8))
    if ipnof not in (5, 7, 8):
        raise NotImplementedError(f"PNOF{ipnof} is not implemented (PNOF5, PNOF7 and GNOF are)")
    method = occ_method or getattr(p, "occ_method", "Softmax")
    if method not in ("Softmax", "Trigonometric"):
        raise NotImplementedError(f"occupation method {method!r} is not implemented")
    return JKEngine.pnof(p.no1, p.ndoc, p.nsoc, p.ncwo, ipnof=ipnof, ista=getattr(p, "ista", 0),
                         occ_method=method, HighSpin=getattr(p, "HighSpin", False), MSpin=getattr(p, "MSpin", 0))


class _Sizes:
    pass


def ocupacion(gamma, no1, ndoc, nalpha, nv, nbf5, ndns, ncwo, HighSpin, occ_method):
    """(n, dn_dgamma) from the parameters (pnof.py:692-701)."""
    if occ_method not in ("Softmax", "Trigonometric"):
        raise NotImplementedError(f"occupation method {occ_method!r} is not implemented")
    p = _Sizes()
    p.no1, p.ndoc, p.nsoc, p.ncwo, p.HighSpin, p.MSpin, p.ipnof = no1, ndoc, ndns - ndoc, ncwo, HighSpin, 0, 5
    pn = _pnof_of(p, occ_method)
    if pn.nbf5 != nbf5 or ncwo * ndoc != nv:
        raise ValueError("nbf5 must equal no1 + ndoc (1 + ncwo) + nsoc and nv = ndoc ncwo")
    r = _plain_engine(p).occupation_step(gamma, pn, outputs=("n", "dn_dgamma"))
    return r["n"], r["dn_dgamma"]


def PNOFi_selector(n, p):
    """(cj12, ck12) for occupations n (pnof.py:657-667)."""
    return _plain_engine(p).pnof_coefficients(n, _pnof_of(p))


def calce(n, cj12, ck12, J_MO, K_MO, H_core, p):
    """Energy (pnof.py:905-990, closed and open shell). As in the reference, the
    diagonals of the caller's cj12 and ck12 are set to zero in place (:913, :917)."""
    for c in (cj12, ck12):
        if isinstance(c, np.ndarray):
            np.fill_diagonal(c, 0)
        else:
            c.fill_diagonal_(0)
    return _plain_engine(p).energy(n, cj12, ck12, J_MO, K_MO, H_core, p.nbeta, p.nalpha, getattr(p, "MSpin", 0))


def calcocce(gamma, J_MO, K_MO, H_core, p):
    """Energy as a function of the occupation parameters (pnof.py:994-1001)."""
    r = _plain_engine(p).occupation_step(gamma, _pnof_of(p), J_MO, K_MO, H_core, outputs=("E",))
    return float(r["E"][0])


def calcoccg(gamma, J_MO, K_MO, H_core, p):
    """Gradient of calcocce with respect to gamma (pnof.py:1004-1025)."""
    return _plain_engine(p).occupation_step(gamma, _pnof_of(p), J_MO, K_MO, H_core, outputs=("grad",))["grad"]
