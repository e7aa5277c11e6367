"""Drop-in replacements for PyNOF's RI JK builds.

``computeJKj(C, I, b_mnl, p)``, ``computeJKH_MO(C, H, I, b_mnl, p)``,
``JKH_MO_tmp(C, H, I, b_mnl, p)``, ``computeLagrange2(n, cj12, ck12, C, H, I,
b_mnl, p)``, ``computeF_RC_driver(J, K, n, H, cj12, ck12, p)``,
``computeF_RO_driver(J, K, n, H, cj12, ck12, p)`` and ``computeLagrange(F, C, p)`` keep
the reference's signatures (src/pynof/integrals.py:172, :307, :665; utils.py:262, :15,
:87, :249), argument meaning and return values:
host FP64 arrays J, K of shape (nbf5, nbf, nbf); J_MO, K_MO (nbf5, nbf5) with
H_core (nbf5,); the orbital-gradient pair H_mat (nbf, nbf5), b_MO
(nbf, nbf5, naux); the Lagrangian elag (nbf, nbf) with Hmat; the generalised Fock
matrices F (nbf5, nbf, nbf). ``I`` is ignored (it is ``[]`` in RI mode,
integrals.py:28-33). Both run on the MI355X through the HIP library; there is no
CPU path. ``p`` must provide ``nbf5``; ``p.RI`` must be true. ``p.precision``
(optional, default "fp64") selects "fp64", "split" or "split_refine".

b_mnl is uploaded to the device on first use and kept there for as long as the
array object lives (the reference uploads it once, integrals.py:51-53).
"""
from __future__ import annotations

import threading
import weakref

from .engine import JKEngine

_cache_lock = threading.Lock()
_engines: dict[int, tuple] = {}


def _engine_for(b_mnl, p) -> JKEngine:
    key = id(b_mnl)
    with _cache_lock:
        hit = _engines.get(key)
        if hit is not None:
            ref, shape, eng = hit
            if ref() is b_mnl and shape == tuple(b_mnl.shape):
                return eng
        eng = JKEngine(int(getattr(p, "device", 0)))
        eng.upload_b(b_mnl, sym_tol=float(getattr(p, "sym_tol", 1e-10)))
        try:
            ref = weakref.ref(b_mnl, lambda _r, k=key: _engines.pop(k, None))
        except TypeError:
            ref = lambda b=b_mnl: b  # noqa: E731 - objects without weakref support stay cached
        _engines[key] = (ref, tuple(b_mnl.shape), eng)
        return eng


def _check_p(p):
    if not getattr(p, "RI", False):
        raise NotImplementedError("pynof_amd implements the density-fitted (p.RI = True) JK build only")
    if not hasattr(p, "nbf5"):
        raise AttributeError("p must provide nbf5")


def computeJKj(C, I, b_mnl, p):
    """J^n, K^n for the first p.nbf5 orbitals (integrals.py:172-191, RI branch)."""
    _check_p(p)
    eng = _engine_for(b_mnl, p)
    r = eng.build(C, p.nbf5, precision=getattr(p, "precision", "fp64"), outputs=("J", "K"))
    return r["J"], r["K"]


def computeJKH_MO(C, H, I, b_mnl, p):
    """J_MO, K_MO, H_core for the first p.nbf5 orbitals (integrals.py:307-326, RI branch)."""
    _check_p(p)
    eng = _engine_for(b_mnl, p)
    r = eng.build(C, p.nbf5, H=H, precision=getattr(p, "precision", "fp64"),
                  outputs=("J_MO", "K_MO", "H_core"))
    return r["J_MO"], r["K_MO"], r["H_core"]


def JKH_MO_tmp(C, H, I, b_mnl, p):
    """(H_mat, b_MO) for the orbital gradient (integrals.py:665-682 with p.RI, i.e.
    Integrals_MO_RI_CPU/_GPU at :703-716): H_mat = C^T H C[:, :nbf5] (nbf, nbf5) and
    b_MO[p, q, P] = sum_mn C[m, p] C[n, q] b[m, n, P] (nbf, nbf5, naux). Host FP64
    arrays, as the reference returns on its CPU path. ``p.precision`` "split"
    gives BF16x3 b_MO; "fp64" and "split_refine" give FP64."""
    _check_p(p)
    eng = _engine_for(b_mnl, p)
    prec = "split" if getattr(p, "precision", "fp64") == "split" else "fp64"
    r = eng.mo_integrals(C, p.nbf5, H=H, precision=prec)
    return r["H_mat"], r["b_MO"]


def computeLagrange2(n, cj12, ck12, C, H, I, b_mnl, p):
    """(elag, Hmat) of the orbital gradient (utils.py:262-382 with p.RI): host FP64
    elag (nbf, nbf) and Hmat (nbf, nbf5). As in the reference, the diagonals of the
    caller's cj12 and ck12 are set to zero in place (:269-270), and for p.MSpin != 0
    elag is returned as zeros (the RI branch fills it for MSpin == 0 only). b_MO is
    formed and contracted on the device and never copied to the host."""
    _check_p(p)
    import numpy as np

    np.fill_diagonal(cj12, 0)
    np.fill_diagonal(ck12, 0)
    eng = _engine_for(b_mnl, p)
    prec = "split" if getattr(p, "precision", "fp64") == "split" else "fp64"
    nbf5 = p.nbf5
    if getattr(p, "MSpin", 0) != 0:
        r = eng.mo_integrals(C, nbf5, H=H, precision=prec, outputs=("H_mat",))
        return np.zeros((eng.nbf, eng.nbf)), r["H_mat"]
    nbeta = int(getattr(p, "nbeta", 0))
    nalpha = int(getattr(p, "nalpha", nbeta))
    r = eng.lagrangian(C, nbf5, H, np.asarray(n)[:nbf5], cj12, ck12, nbeta, nalpha, precision=prec)
    return r["elag"], r["H_mat"]


_plain_engines: dict[int, JKEngine] = {}


def _plain_engine(p) -> JKEngine:
    """An engine with no b, per device, for the Fock-matrix calls."""
    dev = int(getattr(p, "device", 0))
    with _cache_lock:
        eng = _plain_engines.get(dev)
        if eng is None:
            eng = _plain_engines[dev] = JKEngine(dev)
        return eng


def _zero_diagonals(cj12, ck12, p):
    """The reference's np.fill_diagonal(c[ini:, ini:], 0) on the caller's arrays (utils.py:45,
    :50, :121, :138), ini = p.no1 if p.no1 > 1 else 0; torch tensors too."""
    import numpy as np

    ini = p.no1 if p.no1 > 1 else 0
    for c in (cj12, ck12):
        if isinstance(c, np.ndarray):
            np.fill_diagonal(c[ini:, ini:], 0)
        else:
            c[ini:, ini:].fill_diagonal_(0)


def computeF_RC_driver(J, K, n, H, cj12, ck12, p):
    """Closed-shell generalised Fock matrices F (nbf5, nbf, nbf), host FP64
    (computeF_RC_driver, utils.py:15-54; the GPU form :58-84 returns the same host
    array). As in the reference, the diagonals of the caller's cj12[ini:, ini:] and
    ck12[ini:, ini:] are set to zero in place (:45, :50), ini = p.no1 if p.no1 > 1
    else 0. Device (torch) J and K give a device F."""
    _zero_diagonals(cj12, ck12, p)
    nbf5 = int(p.nbf5)
    return _plain_engine(p).fock_rc(J, K, n[:nbf5], H, cj12, ck12, p.no1, p.nbeta, p.nalpha)


def computeF_RO_driver(J, K, n, H, cj12, ck12, p):
    """Restricted open-shell generalised Fock matrices F (nbf5, nbf, nbf), host FP64
    (computeF_RO_driver, utils.py:87-93; CPU form :96-171, GPU form :174-246, both
    returning a host array). The caller's cj12[ini:, ini:] and ck12[ini:, ini:] diagonals
    are zeroed in place as the reference does (:121, :138). Device (torch) J and K give a
    device F."""
    _zero_diagonals(cj12, ck12, p)
    nbf5 = int(p.nbf5)
    return _plain_engine(p).fock_ro(J, K, n[:nbf5], H, cj12, ck12, p.no1, p.nbeta, p.nalpha)


def computeLagrange(F, C, p):
    """Lagrange multipliers elag (nbf, nbf) from the Fock matrices (computeLagrange,
    utils.py:249-259): rows past p.noptorb and columns past p.nbf5 are zero."""
    return _plain_engine(p).lagrange_fock(F, C, p.noptorb)


def computeLagrangeConvergency(elag):
    """(sumdiff, maxdiff) of the Lagrange multipliers (utils.py:405-411), on the device."""
    import numpy as np

    if isinstance(elag, np.ndarray):
        return _plain_engine(_OnDevice(0)).lagrange_convergence(np.ascontiguousarray(elag))
    return _plain_engine(_OnDevice(elag.device.index or 0)).lagrange_convergence(elag)


class _OnDevice:
    """The device field of a param, for the engine cache."""

    def __init__(self, device: int):
        self.device = int(device)


def ENERGY1r(C, n, H, I, b_mnl, cj12, ck12, p):
    """(E, elag, sumdiff, maxdiff) of one orbital step (utils.py:415-432), host outputs as the
    reference returns them. no1 == 0: computeLagrange2 on the device (b_MO never leaves it),
    E = tr elag[:nbf5, :nbf5] + sum n_i Hmat[i, i]. no1 > 0: J^n, K^n, the Fock matrices
    (closed shell for MSpin == 0, restricted open shell otherwise, utils.py:422-425) and elag
    stay on the device (the reference moves J^n, K^n and F across the bus), E as computeE_elec
    (utils.py:385-402) from the build's H_core = diag(C^T H C). As in the reference, the
    diagonals of the caller's cj12 and ck12 are zeroed in place."""
    import numpy as np
    import torch

    _check_p(p)
    n5 = int(p.nbf5)
    n = np.asarray(n, dtype=np.float64)
    if p.no1 == 0:
        elag, Hmat = computeLagrange2(n, cj12, ck12, C, H, I, b_mnl, p)
        E = float(np.trace(elag[:n5, :n5]) + np.dot(n[:n5], np.diagonal(Hmat[:n5, :n5])))
        sumdiff, maxdiff = _engine_for(b_mnl, p).lagrange_convergence(elag)
        return E, elag, sumdiff, maxdiff
    _zero_diagonals(cj12, ck12, p)
    eng = _engine_for(b_mnl, p)
    dev = torch.device("cuda", int(getattr(p, "device", 0)))
    td = lambda x: torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(dev)  # noqa: E731
    Cd, Hd = td(C), td(H)
    prec = getattr(p, "precision", "fp64")
    out = eng.build(Cd, n5, H=Hd, precision=prec, outputs=("J", "K", "H_core"))
    fock = eng.fock_rc if getattr(p, "MSpin", 0) == 0 else eng.fock_ro
    F = fock(out["J"], out["K"], td(n[:n5]), Hd, td(cj12), td(ck12), p.no1, p.nbeta, p.nalpha)
    elag_d = eng.lagrange_fock(F, Cd, p.noptorb)
    sumdiff, maxdiff = eng.lagrange_convergence(elag_d)
    hc = out["H_core"].cpu().numpy()
    b, a = p.nbeta, p.nalpha
    elag = elag_d.cpu().numpy()
    E = float(np.trace(elag[:n5, :n5]) + np.dot(n[:b], hc[:b]) + np.dot(n[a:n5], hc[a:n5]))
    E += 0.5 * float(np.sum(hc[b:a])) if getattr(p, "HighSpin", False) else float(np.dot(n[b:a], hc[b:a]))
    return E, elag, sumdiff, maxdiff


def release_cached_engines() -> None:
    """Drops every cached device copy of b and the per-device engines."""
    with _cache_lock:
        for _, _, eng in _engines.values():
            eng.close()
        _engines.clear()
        for eng in _plain_engines.values():
            eng.close()
        _plain_engines.clear()
