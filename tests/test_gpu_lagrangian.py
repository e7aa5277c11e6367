"""Orbital-gradient Lagrangian (GPU only): pnjk_lagrangian against the FP64 oracle
(computeLagrange2 with RI restated, utils.py:262-355), through the C ABI."""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, _lib, param, synth
from pynof_amd.integrals import computeLagrange2, release_cached_engines

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-10
SPLIT_TOL = 1e-4


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b)))


def inputs(nbf, naux, nbf5, seed):
    b = synth.make_b(nbf, naux, seed, 0.05)
    C = synth.make_C(nbf, seed + 1)
    H = synth.make_H(nbf, seed + 2)
    rng = np.random.default_rng(seed + 3)
    n = rng.uniform(0.02, 0.98, nbf5)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    return b, C, H, n, cj + cj.T, ck + ck.T


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


@pytest.mark.parametrize("nbf,naux,nbf5,nbeta,nalpha", [(1, 1, 1, 1, 1), (6, 7, 4, 2, 2), (24, 116, 24, 5, 5),
                                                        (37, 45, 20, 6, 9), (70, 200, 70, 0, 0),
                                                        (130, 472, 100, 40, 40), (97, 352, 60, 20, 25),
                                                        (420, 96, 400, 100, 110)])
def test_lagrangian_vs_oracle(eng, nbf, naux, nbf5, nbeta, nalpha):
    print(f"case: lagrangian nbf={nbf} naux={naux} nbf5={nbf5} nbeta={nbeta} nalpha={nalpha}", flush=True)
    b, C, H, n, cj, ck = inputs(nbf, naux, nbf5, 600 + nbf)
    eng.upload_b(b)
    el, Hm = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    cj0, ck0 = cj.copy(), ck.copy()
    r = eng.lagrangian(C, nbf5, H, n, cj, ck, nbeta, nalpha)
    assert np.array_equal(cj, cj0) and np.array_equal(ck, ck0)  # the C ABI leaves its inputs alone
    assert r["elag"].shape == (nbf, nbf) and r["H_mat"].shape == (nbf, nbf5)
    assert rel(r["elag"], el) <= FP64_TOL and rel(r["H_mat"], Hm) <= FP64_TOL
    assert np.all(r["elag"][:, nbf5:] == 0.0)
    s = eng.lagrangian(C, nbf5, H, n, cj, ck, nbeta, nalpha, precision="split")
    assert rel(s["elag"], el) <= SPLIT_TOL
    g, go = O.orbital_gradient(r["elag"], nbf5), O.orbital_gradient(el, nbf5)
    assert g.size == 0 or np.max(np.abs(g - go)) <= 1e-9 * max(1.0, np.max(np.abs(go)))


def test_lagrangian_device_io_and_determinism(eng):
    import torch

    nbf, naux, nbf5 = 64, 232, 50
    print(f"case: lagrangian device io nbf={nbf}", flush=True)
    b, C, H, n, cj, ck = inputs(nbf, naux, nbf5, 640)
    eng.upload_b(b)
    el, _ = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, 10, 10)
    dev = [torch.from_numpy(x).cuda() for x in (C, H, n, cj, ck)]
    r = eng.lagrangian(dev[0], nbf5, dev[1], dev[2], dev[3], dev[4], 10, 10)
    torch.cuda.synchronize()
    assert r["elag"].is_cuda and rel(r["elag"].cpu().numpy(), el) <= FP64_TOL
    again = eng.lagrangian(C, nbf5, H, n, cj, ck, 10, 10)
    assert np.array_equal(again["elag"], r["elag"].cpu().numpy())


def test_dropin_compute_lagrange2(gpu_required):
    nbf, naux, nbf5 = 40, 144, 30
    print("case: computeLagrange2 drop-in", flush=True)
    b, C, H, n, cj, ck = inputs(nbf, naux, nbf5, 650)
    p = param(nbf=nbf, nbf5=nbf5, nbfaux=naux, ndoc=12, RI=True)
    el_ref, Hm_ref = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, p.nbeta, p.nalpha)
    cj_in, ck_in = cj.copy(), ck.copy()
    el, Hm = computeLagrange2(n, cj_in, ck_in, C, H, [], b, p)
    assert rel(el, el_ref) <= FP64_TOL and rel(Hm, Hm_ref) <= FP64_TOL
    # the reference zeroes the caller's diagonals in place (utils.py:269-270)
    assert np.all(np.diag(cj_in) == 0.0) and np.all(np.diag(ck_in) == 0.0)
    p.MSpin = 1
    el1, Hm1 = computeLagrange2(n, cj.copy(), ck.copy(), C, H, [], b, p)
    assert np.all(el1 == 0.0) and rel(Hm1, Hm_ref) <= FP64_TOL
    release_cached_engines()


def test_lagrangian_errors(eng):
    nbf, naux, nbf5 = 20, 60, 10
    b, C, H, n, cj, ck = inputs(nbf, naux, nbf5, 660)
    eng.upload_b(b)
    L = _lib.lib()
    el = np.empty((nbf, nbf))
    args = [np.ascontiguousarray(x) for x in (C, H, n, cj, ck)]
    ptr = [a.ctypes.data for a in args]
    ok = L.pnjk_lagrangian(eng._h, ptr[0], nbf, nbf5, ptr[1], nbf, ptr[2], ptr[3], ptr[4], 3, 3, 0,
                           el.ctypes.data, None, None)
    assert ok == _lib.OK
    assert L.pnjk_lagrangian(eng._h, ptr[0], nbf, nbf5, ptr[1], nbf, ptr[2], ptr[3], ptr[4], 3, nbf5 + 1, 0,
                             el.ctypes.data, None, None) == _lib.EINVAL  # nalpha > nbf5
    assert L.pnjk_lagrangian(eng._h, ptr[0], nbf, nbf5, ptr[1], nbf, ptr[2], ptr[3], ptr[4], 4, 3, 0,
                             el.ctypes.data, None, None) == _lib.EINVAL  # nbeta > nalpha
    assert L.pnjk_lagrangian(eng._h, ptr[0], nbf, nbf5, ptr[1], nbf, None, ptr[3], ptr[4], 3, 3, 0,
                             el.ctypes.data, None, None) == _lib.EINVAL  # no occupations
    assert L.pnjk_lagrangian(eng._h, ptr[0], nbf, nbf5, ptr[1], nbf, ptr[2], ptr[3], ptr[4], 3, 3, 2,
                             el.ctypes.data, None, None) == _lib.EUNSUPPORTED
    assert L.pnjk_lagrangian(eng._h, ptr[0], nbf - 1, nbf5, ptr[1], nbf, ptr[2], ptr[3], ptr[4], 3, 3, 0,
                             el.ctypes.data, None, None) == _lib.EINVAL  # ldc < nbf
