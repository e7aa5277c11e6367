"""b from 3-centre integrals on the GPU (pnjk_upload_b_3c) against the oracle's restatement
of integrals.py:35-49, the inverse pnjk_download_b, and the .npy round trip through a
memory map, via the C ABI."""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, io, synth

pytestmark = pytest.mark.gpu


def three_center(nbf, naux, seed):
    rng = np.random.default_rng(seed)
    m = rng.standard_normal((nbf, nbf, naux)) * 0.3
    return 0.5 * (m + m.transpose(1, 0, 2))


def spd_metric(naux, seed, tiny=None):
    rng = np.random.default_rng(seed)
    V, _ = np.linalg.qr(rng.standard_normal((naux, naux)))
    a = np.geomspace(0.05, 20.0, naux)
    if tiny is not None:
        a[tiny] = 20.0 * 1e-16
    return np.ascontiguousarray((V * a) @ V.T)


def rel(a, b):
    return float(np.linalg.norm(a - b) / max(1e-300, np.linalg.norm(b)))


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


@pytest.mark.parametrize("nbf,naux,tiny", [(1, 1, None), (5, 7, None), (24, 116, None), (37, 90, 3), (64, 230, None)])
def test_b_from_3c_vs_oracle(eng, nbf, naux, tiny):
    print(f"case: b from 3c nbf={nbf} naux={naux} tiny={tiny}", flush=True)
    mnk = three_center(nbf, naux, 910 + nbf)
    metric = spd_metric(naux, 911 + naux, tiny)
    ref = O.b_from_3c(mnk, metric)
    eng.upload_b_3c(mnk, metric)
    got = eng.download_b()
    assert rel(got, ref) <= 1e-10
    assert np.array_equal(got, got.transpose(1, 0, 2))


def test_build_from_3c_equals_build_from_b(eng):
    """A JK build after pnjk_upload_b_3c equals one after uploading the oracle's b."""
    import torch

    nbf, naux, nbf5 = 40, 144, 30
    print(f"case: build after b from 3c nbf={nbf}", flush=True)
    mnk = three_center(nbf, naux, 920)
    metric = spd_metric(naux, 921)
    C = synth.make_C(nbf, 922)
    H = synth.make_H(nbf, 923)
    eng.upload_b_3c(torch.from_numpy(mnk).cuda(), torch.from_numpy(metric).cuda())
    a = eng.build(C, nbf5, H=H)
    other = JKEngine(0)
    other.upload_b(O.b_from_3c(mnk, metric))
    b = other.build(C, nbf5, H=H)
    other.close()
    for k in ("J", "K", "J_MO", "K_MO", "H_core"):
        assert rel(a[k], b[k]) <= 1e-10, k
    J, K, J_MO, K_MO, H_core = O.jk_build(C, H, O.b_from_3c(mnk, metric), nbf5)
    assert rel(a["J"], J) <= 1e-10 and rel(a["K"], K) <= 1e-10 and rel(a["K_MO"], K_MO) <= 1e-10


def test_download_and_npy_round_trip(eng, tmp_path):
    import torch

    nbf, naux = 30, 101
    print(f"case: download b and .npy round trip nbf={nbf}", flush=True)
    b = synth.make_b(nbf, naux, 930, 0.05)
    path = tmp_path / "b.npy"
    io.save_b(path, b)
    eng.upload_b(io.load_b(path))  # streamed from the memory map
    assert np.array_equal(eng.download_b(), b)
    dev = torch.empty((nbf, nbf, naux), dtype=torch.float64, device="cuda")
    eng.download_b(dev)
    torch.cuda.synchronize()
    assert np.array_equal(dev.cpu().numpy(), b)
    out = tmp_path / "b_out.npy"
    io.save_b_from_engine(out, eng)
    assert np.array_equal(io.load_b(out), b)


def test_errors(eng):
    from pynof_amd import _lib

    mnk = three_center(4, 5, 940)
    with pytest.raises(ValueError):
        eng.upload_b_3c(mnk, np.eye(4))
    asym = mnk.copy()
    asym[0, 1, 0] += 1.0
    with pytest.raises(_lib.PnjkError):
        eng.upload_b_3c(asym, np.eye(5))
    fresh = JKEngine(0)
    with pytest.raises(_lib.PnjkError):
        fresh.download_b(np.empty((1, 1, 1)))
    fresh.close()
