"""Occupation step on the GPU (pnof.py:155-1025 restated in oracle/pnof_oracle.py):
pnjk_occupation_step, pnjk_pnof_coefficients and pnjk_energy against the FP64
oracle, through the C ABI."""
import ctypes

import numpy as np
import pytest

from oracle import pnof_oracle as O
from pynof_amd import JKEngine, _lib, param, synth
from pynof_amd import pnof as dropin
from pynof_amd.integrals import release_cached_engines

pytestmark = pytest.mark.gpu


def mo_mats(nbf5, seed, naux=12):
    rng = np.random.default_rng(seed)
    B = rng.standard_normal((nbf5, nbf5, naux)) * 0.3
    B = 0.5 * (B + B.transpose(1, 0, 2))
    J = np.einsum("ppP,qqP->pq", B, B)
    K = np.einsum("pqP,pqP->pq", B, B)
    return J, K, rng.standard_normal(nbf5)


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


CONFIGS = [(0, 1, 0, 3), (0, 2, 0, 2), (0, 3, 0, 3), (1, 2, 0, 2), (0, 2, 2, 2), (2, 3, 1, 2), (0, 2, 0, 1),
           (0, 4, 1, 5)]
FUNCS = [(5, 0), (7, 0), (7, 1), (8, 0)]


@pytest.mark.parametrize("method", ["Softmax", "Trigonometric"])
@pytest.mark.parametrize("ipnof,ista", FUNCS)
@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo", CONFIGS)
def test_occupation_step_vs_oracle(eng, method, ipnof, ista, no1, ndoc, nsoc, ncwo):
    print(f"case: occ {method} ipnof={ipnof} ista={ista} cfg={(no1, ndoc, nsoc, ncwo)}", flush=True)
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method)
    J, K, hc = mo_mats(p.nbf5, 100 + 7 * ndoc + ncwo)
    g = np.random.default_rng(200 + ncwo).uniform(-0.8, 1.2, p.nv)
    pn = JKEngine.pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method)
    r = eng.occupation_step(g, pn, J, K, hc, outputs=("E", "grad", "n", "dn_dgamma", "cj12", "ck12"))
    n, dn = O.occupation(g, p)
    cj, ck = O.coefficients(n, p)
    assert np.max(np.abs(r["n"] - n)) <= 1e-14 and np.max(np.abs(r["dn_dgamma"] - dn)) <= 1e-13
    assert np.max(np.abs(r["cj12"] - cj)) <= 1e-14 and np.max(np.abs(r["ck12"] - ck)) <= 1e-14
    E = O.calce(n, cj, ck, J, K, hc, p)
    assert abs(r["E"][0] - E) <= 1e-12 * max(1.0, abs(E))
    grad = O.calcoccg(g, J, K, hc, p)
    assert np.max(np.abs(r["grad"] - grad)) <= 1e-11 * max(1.0, np.max(np.abs(grad)))


@pytest.mark.parametrize("ipnof", [5, 8])
def test_occupation_step_realistic_size(eng, ipnof):
    """nbf5 = 300 (ndoc = 30, ncwo = 9, nv = 270), the size of the large JK workload."""
    print(f"case: occ nbf5=300 ipnof={ipnof}", flush=True)
    p = O.Pnof(0, 30, 0, 9, ipnof=ipnof)
    J, K, hc = mo_mats(p.nbf5, 300, naux=40)
    g = np.random.default_rng(301).uniform(-2.0, 0.5, p.nv)
    pn = JKEngine.pnof(0, 30, 0, 9, ipnof=ipnof)
    r = eng.occupation_step(g, pn, J, K, hc, outputs=("E", "grad"))
    E = O.calcocce(g, J, K, hc, p)
    grad = O.calcoccg(g, J, K, hc, p)
    assert abs(r["E"][0] - E) <= 1e-12 * max(1.0, abs(E))
    assert np.max(np.abs(r["grad"] - grad)) <= 1e-11 * max(1.0, np.max(np.abs(grad)))


def test_device_io_and_determinism(eng):
    import torch

    print("case: occ device io", flush=True)
    p = O.Pnof(1, 3, 1, 3, ipnof=8, occ_method="Trigonometric")
    J, K, hc = mo_mats(p.nbf5, 310)
    g = np.random.default_rng(311).uniform(-0.8, 1.2, p.nv)
    pn = JKEngine.pnof(1, 3, 1, 3, ipnof=8, occ_method="Trigonometric")
    outs = ("E", "grad", "n", "cj12", "ck12")
    host = eng.occupation_step(g, pn, J, K, hc, outputs=outs)
    dev = eng.occupation_step(*(torch.from_numpy(x).cuda() for x in (g,)), pn,
                              *(torch.from_numpy(x).cuda() for x in (J, K, hc)), outputs=outs)
    torch.cuda.synchronize()
    for name in outs:
        assert dev[name].is_cuda and np.array_equal(dev[name].cpu().numpy(), host[name])
    again = eng.occupation_step(g, pn, J, K, hc, outputs=outs)
    for name in outs:
        assert np.array_equal(again[name], host[name])


def test_energy_given_coefficients(eng):
    print("case: pnjk_energy", flush=True)
    p = O.Pnof(2, 3, 1, 2, ipnof=7)
    J, K, hc = mo_mats(p.nbf5, 320)
    rng = np.random.default_rng(321)
    n = rng.uniform(0.0, 1.0, p.nbf5)
    cj = rng.standard_normal((p.nbf5, p.nbf5))  # nonzero diagonals: calce removes them
    ck = rng.standard_normal((p.nbf5, p.nbf5))
    cj0, ck0 = cj.copy(), ck.copy()
    E = eng.energy(n, cj, ck, J, K, hc, p.nbeta, p.nalpha)
    assert np.array_equal(cj, cj0) and np.array_equal(ck, ck0)
    Eo = O.calce(n, cj, ck, J, K, hc, p)
    assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
    cjc, ckc = eng.pnof_coefficients(n, JKEngine.pnof(2, 3, 1, 2, ipnof=7))
    cjo, cko = O.coefficients(n, p)
    assert np.max(np.abs(cjc - cjo)) <= 1e-14 and np.max(np.abs(ckc - cko)) <= 1e-14


def test_chain_from_jk_build(eng):
    """J_MO, K_MO, H_core of a device JK build feed the occupation step on the device."""
    import torch

    nbf, naux, ndoc, ncwo = 40, 144, 6, 4
    nbf5 = ndoc * (1 + ncwo)
    print(f"case: occ chained to a JK build nbf={nbf} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 330, 0.05)
    C = synth.make_C(nbf, 331)
    H = synth.make_H(nbf, 332)
    eng.upload_b(b)
    out = eng.build(torch.from_numpy(C).cuda(), nbf5, H=torch.from_numpy(H).cuda(),
                    outputs=("J_MO", "K_MO", "H_core"))
    pn = JKEngine.pnof(0, ndoc, 0, ncwo, ipnof=8)
    g = torch.from_numpy(np.random.default_rng(333).uniform(-2.0, 0.5, ndoc * ncwo)).cuda()
    r = eng.occupation_step(g, pn, out["J_MO"], out["K_MO"], out["H_core"])
    torch.cuda.synchronize()
    p = O.Pnof(0, ndoc, 0, ncwo, ipnof=8)
    Jh, Kh, hh = (out[k].cpu().numpy() for k in ("J_MO", "K_MO", "H_core"))
    gh = g.cpu().numpy()
    E = O.calcocce(gh, Jh, Kh, hh, p)
    assert abs(r["E"].item() - E) <= 1e-12 * max(1.0, abs(E))
    grad = O.calcoccg(gh, Jh, Kh, hh, p)
    assert np.max(np.abs(r["grad"].cpu().numpy() - grad)) <= 1e-11 * max(1.0, np.max(np.abs(grad)))


def test_dropins(gpu_required):
    print("case: occupation drop-ins", flush=True)
    ndoc, ncwo, no1 = 3, 3, 1
    nbf5 = no1 + ndoc * (1 + ncwo)
    prm = param(nbf=nbf5 + 4, nbf5=nbf5, ndoc=ndoc, no1=no1, ncwo=ncwo, ipnof=7, occ_method="Trigonometric")
    p = O.Pnof(no1, ndoc, 0, ncwo, ipnof=7, occ_method="Trigonometric")
    J, K, hc = mo_mats(nbf5, 340)
    g = np.random.default_rng(341).uniform(-0.8, 1.2, p.nv)
    n, dn = dropin.ocupacion(g, prm.no1, prm.ndoc, prm.nalpha, prm.nv, prm.nbf5, prm.ndns, prm.ncwo,
                             prm.HighSpin, prm.occ_method)
    no, dno = O.occupation(g, p)
    assert np.max(np.abs(n - no)) <= 1e-14 and np.max(np.abs(dn - dno)) <= 1e-13
    cj, ck = dropin.PNOFi_selector(n, prm)
    cjo, cko = O.coefficients(no, p)
    assert np.max(np.abs(cj - cjo)) <= 1e-14 and np.max(np.abs(ck - cko)) <= 1e-14
    E = dropin.calce(n, cj, ck, J, K, hc, prm)
    assert np.all(np.diag(cj) == 0) and np.all(np.diag(ck) == 0)  # zeroed in place, as the reference does
    Eo = O.calce(no, cjo, cko, J, K, hc, p)
    assert abs(E - Eo) <= 1e-12 * max(1.0, abs(Eo))
    assert abs(dropin.calcocce(g, J, K, hc, prm) - Eo) <= 1e-12 * max(1.0, abs(Eo))
    go = O.calcoccg(g, J, K, hc, p)
    assert np.max(np.abs(dropin.calcoccg(g, J, K, hc, prm) - go)) <= 1e-11 * max(1.0, np.max(np.abs(go)))
    prm.ipnof = 4
    with pytest.raises(NotImplementedError):
        dropin.calcocce(g, J, K, hc, prm)
    release_cached_engines()


def test_errors(eng):
    L = _lib.lib()
    J, K, hc = mo_mats(9, 350)
    g = np.zeros(6)
    E = np.zeros(1)

    def call(pn, gamma=g.ctypes.data, Jp=J.ctypes.data, Ep=E.ctypes.data):
        return L.pnjk_occupation_step(eng._h, ctypes.addressof(pn), gamma, Jp, K.ctypes.data, hc.ctypes.data, Ep,
                                      None, None, None, None, None, None)

    assert call(JKEngine.pnof(0, 2, 1, 3)) == _lib.OK
    assert call(JKEngine.pnof(0, 2, 1, 3, MSpin=1)) == _lib.OK  # open shell
    assert call(JKEngine.pnof(0, 2, 1, 3, ipnof=4)) == _lib.EUNSUPPORTED
    bad = JKEngine.pnof(0, 2, 1, 3)
    bad.nbf5 = 10
    assert call(bad) == _lib.EINVAL
    assert call(JKEngine.pnof(0, 2, 1, 3), gamma=None) == _lib.EINVAL
    assert call(JKEngine.pnof(0, 2, 1, 3), Jp=None) == _lib.EINVAL
    assert call(JKEngine.pnof(0, 2, 1, 3), Ep=None) == _lib.EINVAL  # nothing requested
    assert L.pnjk_occupation_step(eng._h, None, None, None, None, None, E.ctypes.data, None, None, None, None, None,
                                  None) == _lib.EINVAL
    assert L.pnjk_energy(eng._h, 9, 4, 3, 0, hc.ctypes.data, J.ctypes.data, K.ctypes.data, J.ctypes.data,
                         K.ctypes.data, hc.ctypes.data, E.ctypes.data, None) == _lib.EINVAL  # nbeta > nalpha
    with pytest.raises(ValueError):
        JKEngine.pnof(0, 2, 1, 3, occ_method="EBI")


OPEN = [(0, 2, 1, 2), (1, 2, 2, 2), (2, 3, 1, 2), (0, 3, 3, 1), (0, 4, 2, 3)]


@pytest.mark.parametrize("method", ["Softmax", "Trigonometric"])
@pytest.mark.parametrize("ipnof,ista", FUNCS)
@pytest.mark.parametrize("no1,ndoc,nsoc,ncwo", OPEN)
def test_open_shell_vs_oracle(eng, method, ipnof, ista, no1, ndoc, nsoc, ncwo):
    """MSpin != 0: coefficients without the closed-shell singly-occupied blocks, the
    open-shell energy (pnof.py:924-988) and gradient (:1027-1098)."""
    print(f"case: occ open shell {method} ipnof={ipnof} ista={ista} cfg={(no1, ndoc, nsoc, ncwo)}", flush=True)
    p = O.Pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method, MSpin=nsoc)
    J, K, hc = mo_mats(p.nbf5, 400 + 7 * ndoc + ncwo)
    g = np.random.default_rng(401 + ncwo).uniform(-0.8, 1.2, p.nv)
    pn = JKEngine.pnof(no1, ndoc, nsoc, ncwo, ipnof=ipnof, ista=ista, occ_method=method, MSpin=nsoc)
    r = eng.occupation_step(g, pn, J, K, hc, outputs=("E", "grad", "cj12", "ck12"))
    n, _ = O.occupation(g, p)
    cj, ck = O.coefficients(n, p)
    assert np.max(np.abs(r["cj12"] - cj)) <= 1e-14 and np.max(np.abs(r["ck12"] - ck)) <= 1e-14
    E = O.calce(n, cj, ck, J, K, hc, p)
    assert abs(r["E"][0] - E) <= 1e-12 * max(1.0, abs(E))
    grad = O.calcoccg(g, J, K, hc, p)
    assert np.max(np.abs(r["grad"] - grad)) <= 1e-11 * max(1.0, np.max(np.abs(grad)))
    Eg = eng.energy(n, cj, ck, J, K, hc, p.nbeta, p.nalpha, MSpin=nsoc)
    assert abs(Eg - E) <= 1e-12 * max(1.0, abs(E))
