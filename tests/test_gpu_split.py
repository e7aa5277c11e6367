"""Split-precision forms against the FP64 oracle (GPU only).

split:        every contraction in BF16x3 (hi*hi + hi*lo + lo*hi, FP32 accumulation).
              Bar: normwise relative error <= 3e-5 for J, K, J_MO, K_MO (a BF16x3
              product carries ~2^-16 relative error); H_core is FP64 (1e-10).
split_refine: J_MO, K_MO in FP64 (the triangular route up to nbf5 = 2 nbf / 3, else the FP64
              form's own transforms, bitwise its J_MO, K_MO; bar 1e-10 against the oracle,
              1e-12 against the FP64 form's), H_core bitwise the FP64 form's, so the energy
              is the FP64 energy to 1e-9 Eh; J, K in BF16x3 (3e-5).
Energies are compared in Eh (FP64_E_EH, SPLIT_E_EH_MID below).
"""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, synth

pytestmark = pytest.mark.gpu
SPLIT_TOL = 3e-5
FP64_TOL = 1e-10
# Energy bars in Eh at nbf=280 (|E| ~ 1.9e4 Eh): the FP64 form against the oracle (FP64 sums
# in another order), and the split form against the FP64 form (measured 1.5e-4..1.2e-3 Eh on
# the MI355X over the serving sizes 24..800; BF16x3 products carry ~2^-16 relative error).
FP64_E_EH = 1e-9
SPLIT_E_EH_MID = 5e-3
NAMES = ("J", "K", "J_MO", "K_MO", "H_core")


def rel(a, r):
    return float(np.max(np.abs(np.asarray(a) - r)) / max(1e-300, np.max(np.abs(r))))


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


SHAPES = [(1, 1, 1), (5, 7, 3), (24, 116, 24), (40, 144, 40), (65, 130, 64), (130, 70, 130), (161, 96, 161),
          (97, 520, 33), (420, 96, 400), (352, 64, 337)]  # the last two: nbf5 > 320 (several M-tiles)


@pytest.mark.parametrize("nbf,naux,nbf5", SHAPES)
def test_split_shapes(eng, nbf, naux, nbf5):
    print(f"case: split shape nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 500 + nbf, 0.05)
    C = synth.make_C(nbf, 600 + nbf)
    H = synth.make_H(nbf, 700 + nbf)
    eng.upload_b(b)
    ref = dict(zip(NAMES, O.jk_build(C, H, b, nbf5)))
    s = eng.build(C, nbf5, H=H, precision="split")
    for k in ("J", "K", "J_MO", "K_MO"):
        assert rel(s[k], ref[k]) <= SPLIT_TOL, (k, rel(s[k], ref[k]))
    assert rel(s["H_core"], ref["H_core"]) <= FP64_TOL
    r = eng.build(C, nbf5, H=H, precision="split_refine")
    for k in ("J_MO", "K_MO", "H_core"):
        assert rel(r[k], ref[k]) <= FP64_TOL, (k, rel(r[k], ref[k]))
    for k in ("J", "K"):
        assert rel(r[k], ref[k]) <= SPLIT_TOL, (k, rel(r[k], ref[k]))


def test_split_energies_mid(eng):
    """Energies at nbf=280, naux=1100 in Eh: the FP64 form against the oracle within
    FP64_E_EH; split_refine's MO outputs (FP64; nbf5 = nbf takes the full route) within 1e-12 of
    the FP64 form's, H_core bitwise, its energy within FP64_E_EH of the FP64 energy; the
    split form within SPLIT_E_EH_MID of the FP64 energy."""
    w = synth.workload(1)
    print(f"case: split energies {w.name}", flush=True)
    b, C, H, n = synth.make_inputs(w)
    eng.upload_b(b)
    ref = O.jk_build(C, H, b, w.nbf5)
    e_ref = O.energy(n, *ref[2:])
    out, res = {}, {}
    for prec in ("fp64", "split", "split_refine"):
        r = eng.build(C, w.nbf5, H=H, precision=prec)
        res[prec] = r
        e = O.energy(n, r["J_MO"], r["K_MO"], r["H_core"])
        out[prec] = (e, {k: rel(r[k], a) for k, a in zip(NAMES, ref)})
        print(prec, "E", e, "E - E_oracle (Eh)", e - e_ref, out[prec][1], flush=True)
    assert abs(out["fp64"][0] - e_ref) <= FP64_E_EH
    for k in ("J_MO", "K_MO"):
        assert rel(res["split_refine"][k], res["fp64"][k]) <= 1e-12, k
    assert np.array_equal(res["split_refine"]["H_core"], res["fp64"]["H_core"])
    assert abs(out["split_refine"][0] - out["fp64"][0]) <= FP64_E_EH
    assert abs(out["split"][0] - out["fp64"][0]) <= SPLIT_E_EH_MID
    for k in ("J", "K", "J_MO", "K_MO"):
        assert out["split"][1][k] <= SPLIT_TOL


def test_split_is_deterministic(eng):
    print("case: split determinism nbf=90 naux=200 nbf5=60", flush=True)
    b = synth.make_b(90, 200, 9, 0.05)
    C = synth.make_C(90, 10)
    eng.upload_b(b)
    a = eng.build(C, 60, precision="split", outputs=("J", "K", "J_MO", "K_MO"))
    c = eng.build(C, 60, precision="split", outputs=("J", "K", "J_MO", "K_MO"))
    for k in a:
        assert np.array_equal(a[k], c[k]), k


@pytest.mark.parametrize("nbf,naux,nbf5", [(2, 3, 1), (5, 7, 3), (40, 100, 39), (60, 200, 40), (97, 520, 33),
                                          (130, 400, 70), (352, 64, 337), (40, 144, 40), (10, 40, 4),
                                          (33, 100, 10), (100, 300, 20), (150, 400, 50), (420, 96, 160)])
def test_split_mo_only_builds(eng, nbf, naux, nbf5):
    """A split build asked for J_MO, K_MO and H_core only takes the BF16x3 triangular route
    when nbf5 <= 2 nbf / 5, where it is the faster route (T16 = L C per mu with K = mu + 1,
    W = C^T T16 in FP32, K_MO, d and J_MO reduced in FP64): J_MO, K_MO against the oracle at
    the split bar and within it of the full split build's, H_core bitwise the full build's;
    past the threshold the full route is kept (bitwise); host and device inputs give the
    same bits."""
    import torch

    print(f"case: split MO-only nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 71, 0.05)
    C = synth.make_C(nbf, 72)
    H = synth.make_H(nbf, 73)
    eng.upload_b(b)
    full = eng.build(C, nbf5, H=H, precision="split")
    mo = eng.build(C, nbf5, H=H, precision="split", outputs=("J_MO", "K_MO", "H_core"))
    ref = dict(zip(NAMES, O.jk_build(C, H, b, nbf5)))
    for k in ("J_MO", "K_MO"):
        assert rel(mo[k], ref[k]) <= SPLIT_TOL, (k, rel(mo[k], ref[k]))
        assert rel(mo[k], full[k]) <= SPLIT_TOL, (k, rel(mo[k], full[k]))
        assert np.array_equal(mo[k], mo[k].T), k
    assert np.array_equal(mo["H_core"], full["H_core"])
    if 5 * nbf5 > 2 * nbf:  # the full route
        for k in ("J_MO", "K_MO"):
            assert np.array_equal(mo[k], full[k]), k
    else:  # the route's reduction forms both diagonals in one order: d_p . d_p = sum Y_pp^2
        assert np.array_equal(np.diag(mo["J_MO"]), np.diag(mo["K_MO"]))
    dev = eng.build(torch.from_numpy(C).cuda(), nbf5, H=torch.from_numpy(H).cuda(), precision="split",
                    outputs=("J_MO", "K_MO", "H_core"))
    torch.cuda.synchronize()
    for k in mo:
        assert np.array_equal(dev[k].cpu().numpy(), mo[k]), k


@pytest.mark.parametrize("nbf,naux,nbf5", [(40, 144, 40), (130, 70, 130), (97, 300, 80), (161, 96, 161), (5, 7, 4)])
def test_refine_full_route_is_fp64_bitwise(eng, nbf, naux, nbf5):
    """Past nbf5 = 2 nbf / 3 split_refine forms J_MO and K_MO with the FP64 form's half- and
    quarter transforms: bitwise the FP64 form's (H_core too); its J^n and K^n, from the split
    planes of the FP64 d and X, within the split bar of the FP64 form's."""
    assert 3 * nbf5 > 2 * nbf
    print(f"case: refine full route nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 81, 0.05)
    C = synth.make_C(nbf, 82)
    H = synth.make_H(nbf, 83)
    eng.upload_b(b)
    f = eng.build(C, nbf5, H=H, precision="fp64")
    r = eng.build(C, nbf5, H=H, precision="split_refine")
    for k in ("J_MO", "K_MO", "H_core"):
        assert np.array_equal(r[k], f[k]), k
    for k in ("J", "K"):
        assert rel(r[k], f[k]) <= SPLIT_TOL, (k, rel(r[k], f[k]))
