"""Generalised Fock matrices and the Lagrangian from them (GPU only): pnjk_fock_rc,
pnjk_fock_ro and pnjk_lagrange_fock against the FP64 oracle (computeF_RC_CPU,
utils.py:25-54, computeF_RO_CPU, :96-171, and computeLagrange, :249-259, restated), through
the C ABI."""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, _lib, param, synth
from pynof_amd.integrals import computeF_RC_driver, computeF_RO_driver, computeLagrange, release_cached_engines

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-10


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b)))


def fock_inputs(nbf, nbf5, seed):
    rng = np.random.default_rng(seed)
    J = rng.standard_normal((nbf5, nbf, nbf))
    K = rng.standard_normal((nbf5, nbf, nbf))
    J = J + J.transpose(0, 2, 1)
    K = K + K.transpose(0, 2, 1)
    H = synth.make_H(nbf, seed + 1)
    n = rng.uniform(0.02, 0.98, nbf5)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    return J, K, H, n, cj + cj.T, ck + ck.T


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


# (nbf, nbf5, no1, nbeta, nalpha): nbf^2 odd (staged J, K), ini > 0, open shells, and
# (260, 90) for the large GEMM tile.
CASES = [(1, 1, 0, 1, 1), (5, 3, 0, 1, 2), (24, 24, 2, 8, 8), (37, 20, 3, 6, 9), (64, 50, 0, 10, 10),
         (130, 100, 0, 40, 40), (260, 90, 4, 30, 34)]


@pytest.mark.parametrize("nbf,nbf5,no1,nbeta,nalpha", CASES)
def test_fock_vs_oracle(eng, nbf, nbf5, no1, nbeta, nalpha):
    print(f"case: fock nbf={nbf} nbf5={nbf5} no1={no1} nbeta={nbeta} nalpha={nalpha}", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 700 + nbf)
    F_ref = O.fock_rc(J, K, n, H, cj, ck, nbeta, nalpha, no1)
    cj0, ck0 = cj.copy(), ck.copy()
    F = eng.fock_rc(J, K, n, H, cj, ck, no1, nbeta, nalpha)
    assert np.array_equal(cj, cj0) and np.array_equal(ck, ck0)  # the C ABI leaves its inputs alone
    assert F.shape == (nbf5, nbf, nbf) and rel(F, F_ref) <= FP64_TOL
    C = synth.make_C(nbf, 710 + nbf)
    for noptorb in sorted({nbf, max(nbf5, nbf - 3)}):
        el_ref = O.lagrange_fock(F_ref, C, noptorb)
        el = eng.lagrange_fock(F, C, noptorb)
        assert el.shape == (nbf, nbf) and rel(el, el_ref) <= FP64_TOL
        assert np.all(el[noptorb:] == 0.0) and np.all(el[:, nbf5:] == 0.0)


# (nbf, nbf5, no1, nbeta, nalpha) with singly occupied orbitals [nbeta, nalpha); nsoc = 0 too.
RO_CASES = [(1, 1, 0, 0, 1), (5, 3, 0, 1, 2), (24, 24, 2, 8, 10), (37, 20, 3, 6, 9), (64, 50, 0, 10, 13),
            (130, 100, 1, 40, 45), (260, 90, 4, 30, 34), (40, 30, 2, 12, 12)]


@pytest.mark.parametrize("nbf,nbf5,no1,nbeta,nalpha", RO_CASES)
def test_fock_ro_vs_oracle(eng, nbf, nbf5, no1, nbeta, nalpha):
    print(f"case: fock_ro nbf={nbf} nbf5={nbf5} no1={no1} nbeta={nbeta} nalpha={nalpha}", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 800 + nbf)
    n[nbeta:nalpha] = 0.5
    F_ref = O.fock_ro(J, K, n, H, cj, ck, nbeta, nalpha, no1)
    cj0, ck0 = cj.copy(), ck.copy()
    F = eng.fock_ro(J, K, n, H, cj, ck, no1, nbeta, nalpha)
    assert np.array_equal(cj, cj0) and np.array_equal(ck, ck0)
    assert F.shape == (nbf5, nbf, nbf) and rel(F, F_ref) <= FP64_TOL
    if nbeta == nalpha:  # no singly occupied orbitals: the closed-shell matrices
        assert rel(F, eng.fock_rc(J, K, n, H, cj, ck, no1, nbeta, nalpha)) <= 1e-14
    C = synth.make_C(nbf, 810 + nbf)
    el = eng.lagrange_fock(F, C, nbf)
    assert rel(el, O.lagrange_fock(F_ref, C, nbf)) <= FP64_TOL


def test_fock_ro_device_io(eng):
    import torch

    nbf, nbf5, no1, nbeta, nalpha = 48, 40, 2, 12, 15
    print(f"case: fock_ro device io nbf={nbf} nbf5={nbf5}", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 820)
    F_ref = O.fock_ro(J, K, n, H, cj, ck, nbeta, nalpha, no1)
    dev = [torch.from_numpy(x).cuda() for x in (J, K, n, H, cj, ck)]
    F = eng.fock_ro(*dev, no1, nbeta, nalpha)
    torch.cuda.synchronize()
    assert F.is_cuda and rel(F.cpu().numpy(), F_ref) <= FP64_TOL
    F2 = eng.fock_ro(J, K, n, H, cj, ck, no1, nbeta, nalpha)  # host in, host out
    assert np.array_equal(F2, F.cpu().numpy())


def test_fock_device_io_staging_and_determinism(eng):
    import torch

    nbf, nbf5 = 48, 40
    print(f"case: fock device io nbf={nbf} nbf5={nbf5}", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 760)
    F_ref = O.fock_rc(J, K, n, H, cj, ck, 12, 15, 0)
    dev = [torch.from_numpy(x).cuda() for x in (J, K, n, H, cj, ck)]
    F = eng.fock_rc(*dev, 0, 12, 15)
    torch.cuda.synchronize()
    assert F.is_cuda and rel(F.cpu().numpy(), F_ref) <= FP64_TOL
    # J, K at an 8-byte offset (not 16-byte aligned) take the staged path: same bits.
    size = J.size
    raw = torch.empty(2 * size + 2, dtype=torch.float64, device="cuda")
    Jm = raw[1:1 + size].view(J.shape)
    Km = raw[size + 1:2 * size + 1].view(K.shape)
    Jm.copy_(dev[0])
    Km.copy_(dev[1])
    assert Jm.data_ptr() % 16 == 8
    F2 = eng.fock_rc(Jm, Km, *dev[2:], 0, 12, 15)
    torch.cuda.synchronize()
    assert torch.equal(F, F2)
    F3 = eng.fock_rc(J, K, n, H, cj, ck, 0, 12, 15)  # host in, host out
    assert np.array_equal(F3, F.cpu().numpy())
    C = synth.make_C(nbf, 761)
    el = eng.lagrange_fock(F, torch.from_numpy(C).cuda(), nbf)
    torch.cuda.synchronize()
    el_h = eng.lagrange_fock(F3, C, nbf)
    assert el.is_cuda and np.array_equal(el.cpu().numpy(), el_h)
    assert rel(el_h, O.lagrange_fock(F_ref, C, nbf)) <= FP64_TOL


@pytest.mark.parametrize("nbf,naux,nbf5,nbeta,nalpha", [(24, 116, 24, 5, 5), (37, 45, 20, 6, 9)])
def test_fock_route_equals_lagrangian(eng, nbf, naux, nbf5, nbeta, nalpha):
    """With ini = 0 and noptorb = nbf, J^n, K^n -> F -> elag (all on the device) gives the
    multipliers pnjk_lagrangian forms from b_MO (utils.py:415-427 takes either route)."""
    import torch

    print(f"case: fock route vs lagrangian nbf={nbf} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 780 + nbf, 0.05)
    C = synth.make_C(nbf, 781 + nbf)
    H = synth.make_H(nbf, 782 + nbf)
    rng = np.random.default_rng(783 + nbf)
    n = rng.uniform(0.02, 0.98, nbf5)
    cj = rng.standard_normal((nbf5, nbf5))
    ck = rng.standard_normal((nbf5, nbf5))
    cj, ck = cj + cj.T, ck + ck.T
    eng.upload_b(b)
    Cd, Hd = torch.from_numpy(C).cuda(), torch.from_numpy(H).cuda()
    jk = eng.build(Cd, nbf5, outputs=("J", "K"))
    F = eng.fock_rc(jk["J"], jk["K"], torch.from_numpy(n).cuda(), Hd, torch.from_numpy(cj).cuda(),
                    torch.from_numpy(ck).cuda(), 0, nbeta, nalpha)
    el_f = eng.lagrange_fock(F, Cd, nbf).cpu().numpy()
    el_l = eng.lagrangian(C, nbf5, H, n, cj, ck, nbeta, nalpha)["elag"]
    assert rel(el_f, el_l) <= FP64_TOL
    el_ref, _ = O.lagrange2_ri(n, cj, ck, C, H, b, nbf5, nbeta, nalpha)
    assert rel(el_f, el_ref) <= FP64_TOL


def test_dropin_fock_and_lagrange(gpu_required):
    nbf, nbf5, no1 = 30, 22, 3
    print("case: computeF_RC_driver / computeLagrange drop-ins", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 790)
    p = param(nbf=nbf, nbf5=nbf5, ndoc=9, RI=True, no1=no1, noptorb=nbf - 2)
    F_ref = O.fock_rc(J, K, n, H, cj, ck, p.nbeta, p.nalpha, no1)
    cj_in, ck_in = cj.copy(), ck.copy()
    F = computeF_RC_driver(J, K, n, H, cj_in, ck_in, p)
    assert isinstance(F, np.ndarray) and rel(F, F_ref) <= FP64_TOL
    # the reference zeroes the caller's [ini:, ini:] diagonals in place (utils.py:45, :50)
    assert np.all(np.diag(cj_in)[no1:] == 0.0) and np.array_equal(np.diag(cj_in)[:no1], np.diag(cj)[:no1])
    assert np.all(np.diag(ck_in)[no1:] == 0.0) and np.array_equal(np.diag(ck_in)[:no1], np.diag(ck)[:no1])
    C = synth.make_C(nbf, 791)
    el = computeLagrange(F, C, p)
    assert rel(el, O.lagrange_fock(F_ref, C, p.noptorb)) <= FP64_TOL
    release_cached_engines()


def test_dropin_fock_ro(gpu_required):
    nbf, nbf5, no1 = 30, 22, 3
    print("case: computeF_RO_driver drop-in", flush=True)
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 830)
    p = param(nbf=nbf, nbf5=nbf5, ndoc=7, nsoc=2, RI=True, no1=no1, noptorb=nbf - 2)
    F_ref = O.fock_ro(J, K, n, H, cj, ck, p.nbeta, p.nalpha, no1)
    cj_in, ck_in = cj.copy(), ck.copy()
    F = computeF_RO_driver(J, K, n, H, cj_in, ck_in, p)
    assert isinstance(F, np.ndarray) and rel(F, F_ref) <= FP64_TOL
    # the reference zeroes the caller's [ini:, ini:] diagonals in place (utils.py:121, :138)
    assert np.all(np.diag(cj_in)[no1:] == 0.0) and np.array_equal(np.diag(cj_in)[:no1], np.diag(cj)[:no1])
    assert np.all(np.diag(ck_in)[no1:] == 0.0) and np.array_equal(np.diag(ck_in)[:no1], np.diag(ck)[:no1])
    release_cached_engines()


def test_fock_errors(eng):
    nbf, nbf5 = 12, 8
    J, K, H, n, cj, ck = fock_inputs(nbf, nbf5, 795)
    L = _lib.lib()
    F = np.empty((nbf5, nbf, nbf))
    el = np.empty((nbf, nbf))
    C = synth.make_C(nbf, 796)
    pj, pk, ph, pn, pcj, pck = (np.ascontiguousarray(x).ctypes.data for x in (J, K, H, n, cj, ck))
    call = lambda nb5=nbf5, ldh=nbf, no1=0, nbeta=3, nalpha=3, Fp=F.ctypes.data: L.pnjk_fock_rc(  # noqa: E731
        eng._h, pj, pk, nbf, nb5, pn, ph, ldh, pcj, pck, no1, nbeta, nalpha, Fp, None)
    assert call() == _lib.OK
    assert call(nalpha=nbf5 + 1) == _lib.EINVAL
    assert call(nbeta=4) == _lib.EINVAL  # nbeta > nalpha
    assert call(ldh=nbf - 1) == _lib.EINVAL
    assert call(nb5=nbf + 1) == _lib.EINVAL
    assert call(no1=-1) == _lib.EINVAL
    assert call(Fp=None) == _lib.EINVAL
    assert L.pnjk_lagrange_fock(eng._h, F.ctypes.data, nbf, nbf5, C.ctypes.data, nbf, nbf, el.ctypes.data,
                                None) == _lib.OK
    assert L.pnjk_lagrange_fock(eng._h, F.ctypes.data, nbf, nbf5, C.ctypes.data, nbf - 1, nbf, el.ctypes.data,
                                None) == _lib.EINVAL  # ldc < nbf
    assert L.pnjk_lagrange_fock(eng._h, F.ctypes.data, nbf, nbf5, C.ctypes.data, nbf, nbf + 1, el.ctypes.data,
                                None) == _lib.EINVAL  # noptorb > nbf
    with pytest.raises(ValueError):
        eng.fock_rc(J, K[:-1], n, H, cj, ck, 0, 3, 3)
    ro = lambda nbeta=3, nalpha=5, Fp=F.ctypes.data: L.pnjk_fock_ro(  # noqa: E731
        eng._h, pj, pk, nbf, nbf5, pn, ph, nbf, pcj, pck, 0, nbeta, nalpha, Fp, None)
    assert ro() == _lib.OK
    assert ro(nalpha=nbf5 + 1) == _lib.EINVAL
    assert ro(nbeta=6) == _lib.EINVAL
    assert ro(Fp=None) == _lib.EINVAL
    assert L.pnjk_fock_ro(None, pj, pk, nbf, nbf5, pn, ph, nbf, pcj, pck, 0, 3, 5, F.ctypes.data,
                          None) == _lib.EINVAL
    assert "pnjk_fock_ro" in L.pnjk_last_error().decode()
    with pytest.raises(ValueError):
        eng.fock_ro(J, K, n[:-1], H, cj, ck, 0, 3, 3)
