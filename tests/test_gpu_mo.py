"""Orbital-gradient MO integrals (GPU only): pnjk_mo_integrals against the FP64
oracle (Integrals_MO_RI_CPU restated, integrals.py:703-707), through the C ABI."""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, _lib, param, synth
from pynof_amd.integrals import JKH_MO_tmp, release_cached_engines

pytestmark = pytest.mark.gpu
FP64_TOL = 1e-10
SPLIT_TOL = 3e-5
SHAPES = [(1, 1, 1), (5, 7, 3), (24, 116, 24), (37, 45, 20), (70, 200, 70), (130, 472, 100), (97, 352, 60),
          (420, 96, 400)]  # the last: nbf5 > 320


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b)))


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


@pytest.mark.parametrize("nbf,naux,nbf5", SHAPES)
def test_mo_integrals_vs_oracle(eng, nbf, naux, nbf5):
    print(f"case: mo integrals nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 500 + nbf, 0.05)
    C = synth.make_C(nbf, 501 + nbf)
    H = synth.make_H(nbf, 502 + nbf)
    eng.upload_b(b)
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    r = eng.mo_integrals(C, nbf5, H=H, precision="fp64")
    assert r["b_MO"].shape == (nbf, nbf5, naux) and r["H_mat"].shape == (nbf, nbf5)
    assert rel(r["b_MO"], bm) <= FP64_TOL and rel(r["H_mat"], Hm) <= FP64_TOL
    s = eng.mo_integrals(C, nbf5, H=H, precision="split")
    assert rel(s["b_MO"], bm) <= SPLIT_TOL
    assert rel(s["H_mat"], Hm) <= FP64_TOL  # H_mat is always FP64


def test_mo_integrals_device_io_and_single_outputs(eng):
    import torch

    nbf, naux, nbf5 = 64, 232, 50
    print(f"case: mo integrals device io nbf={nbf}", flush=True)
    b = synth.make_b(nbf, naux, 510, 0.05)
    C = synth.make_C(nbf, 511)
    H = synth.make_H(nbf, 512)
    eng.upload_b(b)
    Hm, bm = O.integrals_mo_ri(C, H, b, nbf5)
    Cd, Hd = torch.from_numpy(C).cuda(), torch.from_numpy(H).cuda()
    r = eng.mo_integrals(Cd, nbf5, H=Hd)
    torch.cuda.synchronize()
    assert r["b_MO"].is_cuda and rel(r["b_MO"].cpu().numpy(), bm) <= FP64_TOL
    assert rel(r["H_mat"].cpu().numpy(), Hm) <= FP64_TOL
    only_b = eng.mo_integrals(C, nbf5, outputs=("b_MO",))
    assert set(only_b) == {"b_MO"} and rel(only_b["b_MO"], bm) <= FP64_TOL
    only_h = eng.mo_integrals(C, nbf5, H=H, outputs=("H_mat",))
    assert set(only_h) == {"H_mat"} and rel(only_h["H_mat"], Hm) <= FP64_TOL
    # F-order C (from eigh, energy.py:57) gives the same result
    rf = eng.mo_integrals(np.asfortranarray(C), nbf5, H=H)
    assert np.array_equal(rf["b_MO"], eng.mo_integrals(C, nbf5, H=H)["b_MO"])


def test_mo_block_matches_build(eng):
    """The nbf5 x nbf5 block of b_MO reproduces the JK build's MO outputs, and the
    split b_MO's diagonal block agrees with the split build's J_MO."""
    w = synth.workload(1)
    print(f"case: mo block vs build {w.name}", flush=True)
    b, C, H, n = synth.make_inputs(w)
    eng.upload_b(b)
    r = eng.mo_integrals(C, w.nbf5, H=H)
    blk = r["b_MO"][: w.nbf5]
    out = eng.build(C, w.nbf5, H=H, outputs=("J_MO", "K_MO", "H_core"))
    dg = np.einsum("ppP->pP", blk)
    assert rel(dg @ dg.T, out["J_MO"]) <= 1e-10
    assert rel(np.einsum("pqP,pqP->pq", blk, blk), out["K_MO"]) <= 1e-10
    assert rel(np.diag(r["H_mat"][: w.nbf5]), out["H_core"]) <= 1e-12
    assert rel(blk, blk.transpose(1, 0, 2)) <= 1e-12


def test_dropin_jkh_mo_tmp(gpu_required):
    nbf, naux, nbf5 = 40, 144, 30
    print("case: JKH_MO_tmp drop-in", flush=True)
    b = synth.make_b(nbf, naux, 520, 0.05)
    C = synth.make_C(nbf, 521)
    H = synth.make_H(nbf, 522)
    p = param(nbf=nbf, nbf5=nbf5, nbfaux=naux, RI=True)
    Hm, bm = JKH_MO_tmp(C, H, [], b, p)
    Hr, br = O.integrals_mo_ri(C, H, b, nbf5)
    assert isinstance(bm, np.ndarray) and bm.shape == (nbf, nbf5, naux)
    assert rel(bm, br) <= FP64_TOL and rel(Hm, Hr) <= FP64_TOL
    release_cached_engines()


def test_mo_integrals_errors(eng):
    nbf, naux, nbf5 = 20, 60, 10
    b = synth.make_b(nbf, naux, 530, 0.05)
    C = synth.make_C(nbf, 531)
    H = synth.make_H(nbf, 532)
    eng.upload_b(b)
    L = _lib.lib()
    out = np.empty((nbf, nbf5, naux))
    Cn = np.ascontiguousarray(C)
    # fewer columns than nbf: every column of C is used
    assert L.pnjk_mo_integrals(eng._h, Cn.ctypes.data, nbf - 1, nbf5, None, 0, 0, out.ctypes.data, None,
                               None) == _lib.EINVAL
    assert L.pnjk_mo_integrals(eng._h, Cn.ctypes.data, nbf, nbf5, None, 0, 2, out.ctypes.data, None,
                               None) == _lib.EUNSUPPORTED
    hm = np.empty((nbf, nbf5))
    assert L.pnjk_mo_integrals(eng._h, Cn.ctypes.data, nbf, nbf5, None, 0, 0, None, hm.ctypes.data,
                               None) == _lib.EINVAL  # H_mat without H
    assert L.pnjk_mo_integrals(eng._h, Cn.ctypes.data, nbf, nbf5, None, 0, 0, None, None, None) == _lib.EINVAL
    assert L.pnjk_mo_integrals(eng._h, Cn.ctypes.data, nbf, nbf + 1, None, 0, 0, out.ctypes.data, None,
                               None) == _lib.EINVAL
    with pytest.raises(ValueError):
        eng.mo_integrals(C[:, :nbf5], nbf5)
    fresh = JKEngine(0)
    assert L.pnjk_mo_integrals(fresh._h, Cn.ctypes.data, nbf, nbf5, None, 0, 0, out.ctypes.data, None,
                               None) == _lib.ESTATE
    fresh.close()
