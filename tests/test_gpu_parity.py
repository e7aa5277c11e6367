"""FP64 parity of the HIP JK build against the FP64 oracle (GPU only).

Bar: normwise relative error max|gpu - oracle| / max|oracle| <= 1e-10 for every
output of the FP64 form (the stated FP64 parity tolerance), energy within
1e-10 relative.
"""
import json
import os

import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, PnjkError, computeJKH_MO, computeJKj, param, synth
from pynof_amd import _lib

pytestmark = pytest.mark.gpu
TOL = 1e-10
HERE = os.path.dirname(os.path.abspath(__file__))
NAMES = ("J", "K", "J_MO", "K_MO", "H_core")


def rel(a, r):
    return float(np.max(np.abs(np.asarray(a) - r)) / max(1e-300, np.max(np.abs(r))))


@pytest.fixture(scope="module")
def eng(gpu_required):
    e = JKEngine(0)
    yield e
    e.close()


def _say(msg):
    print(msg, flush=True)


def test_water_golden(eng):
    z = np.load(os.path.join(HERE, "golden", "water_fp64.npz"))
    meta = json.load(open(os.path.join(HERE, "golden", "water_fp64.json")))
    _say("case: water golden nbf=24 naux=116 nbf5=24")
    eng.upload_b(z["b"])
    r = eng.build(z["C"], meta["nbf5"], H=z["H"])
    for k in NAMES:
        assert rel(r[k], z[k]) <= TOL, (k, rel(r[k], z[k]))
    e = O.energy(z["n"], r["J_MO"], r["K_MO"], r["H_core"])
    assert abs(e - meta["E"]) <= TOL * abs(meta["E"])


# (the last two: nbf5 > 320, so three 160-row M-tiles in the half-transform and J^n and a
# quarter-transform window of two M-tiles; aux slices keep the oracle quick)
SHAPES = [(1, 1, 1), (2, 3, 1), (5, 7, 3), (17, 33, 9), (40, 144, 40), (65, 130, 64), (129, 200, 100),
          (130, 70, 130), (161, 96, 161), (97, 520, 33), (420, 96, 400), (352, 64, 337)]


@pytest.mark.parametrize("nbf,naux,nbf5", SHAPES)
def test_shapes_vs_oracle(eng, nbf, naux, nbf5):
    _say(f"case: shape nbf={nbf} naux={naux} nbf5={nbf5}")
    b = synth.make_b(nbf, naux, 100 + nbf, 0.05)
    C = synth.make_C(nbf, 200 + nbf)
    H = synth.make_H(nbf, 300 + nbf)
    eng.upload_b(b)
    r = eng.build(C, nbf5, H=H)
    ref = O.jk_build(C, H, b, nbf5)
    for k, a in zip(NAMES, ref):
        assert rel(r[k], a) <= TOL, (k, rel(r[k], a))


def test_mid_workload(eng):
    w = synth.workload(1)
    _say(f"case: {w.name}")
    b, C, H, n = synth.make_inputs(w)
    eng.upload_b(b)
    r = eng.build(C, w.nbf5, H=H)
    ref = O.jk_build(C, H, b, w.nbf5)
    for k, a in zip(NAMES, ref):
        assert rel(r[k], a) <= TOL, (k, rel(r[k], a))
    e_ref = O.energy(n, *ref[2:])
    e = O.energy(n, r["J_MO"], r["K_MO"], r["H_core"])
    assert abs(e - e_ref) <= TOL * abs(e_ref)


def test_device_inputs_and_outputs(eng):
    import torch

    _say("case: device inputs/outputs nbf=70 naux=150 nbf5=50")
    nbf, naux, nbf5 = 70, 150, 50
    b = synth.make_b(nbf, naux, 7, 0.05)
    C = synth.make_C(nbf, 8)
    H = synth.make_H(nbf, 9)
    eng.upload_b(b)
    host = eng.build(C, nbf5, H=H)
    e2 = JKEngine(0)
    e2.upload_b(torch.from_numpy(b).cuda())
    dev = e2.build(torch.from_numpy(C).cuda(), nbf5, H=torch.from_numpy(H).cuda())
    torch.cuda.synchronize()
    for k in NAMES:
        assert dev[k].is_cuda
        assert np.array_equal(dev[k].cpu().numpy(), host[k]), k
    e2.close()


def test_fortran_order_C(eng):
    _say("case: F-order C nbf=30 naux=64 nbf5=20")
    nbf, naux, nbf5 = 30, 64, 20
    b = synth.make_b(nbf, naux, 17, 0.05)
    C = np.asfortranarray(synth.make_C(nbf, 18))
    eng.upload_b(b)
    r = eng.build(C, nbf5, outputs=("J", "K"))
    J, K = O.jkj_ri(C, b, nbf5)
    assert rel(r["J"], J) <= TOL and rel(r["K"], K) <= TOL


def test_dropin_matches_reference_sequence(gpu_required):
    _say("case: drop-in computeJKj / computeJKH_MO water")
    w = synth.workload(0)
    b, C, H, n = synth.make_inputs(w)
    p = param(w.nbf, w.nbf5, nbfaux=w.naux, ndoc=w.ndoc)
    J, K = computeJKj(C, [], b, p)
    JM, KM, Hc = computeJKH_MO(C, H, [], b, p)
    Jr, Kr = O.jkj_ri(C, b, w.nbf5)
    JMr, KMr, Hr = O.jkh_mo_ri(C, H, b, w.nbf5)
    for a, r in ((J, Jr), (K, Kr), (JM, JMr), (KM, KMr), (Hc, Hr)):
        assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == r.shape
        assert rel(a, r) <= TOL
    # the cached device copy of b is reused for a rotated C
    C2 = C @ synth.rotation(w.nbf, 6000)
    J2, _ = computeJKj(C2, [], b, p)
    assert rel(J2, O.jkj_ri(C2, b, w.nbf5)[0]) <= TOL


def test_errors(eng):
    _say("case: error paths")
    fresh = JKEngine(0)
    with pytest.raises(PnjkError) as ei:
        fresh.build(np.eye(3), 2)
    assert ei.value.code == _lib.ESTATE
    b = synth.make_b(6, 5, 1, 0.05)
    fresh.upload_b(b)
    with pytest.raises(ValueError):
        fresh.build(np.eye(6), 7)
    with pytest.raises(PnjkError) as ei:
        fresh.build(np.eye(6), 3, precision=7, outputs=("J", "K"))
    assert ei.value.code in (_lib.EINVAL, _lib.EUNSUPPORTED)
    bad = b.copy()
    bad[1, 2, 3] = np.nan
    with pytest.raises(PnjkError) as ei:
        fresh.upload_b(bad)
    assert ei.value.code == _lib.EINVAL and "non-finite" in str(ei.value)
    asym = b.copy()
    asym[4, 1, 0] += 1e-3
    with pytest.raises(PnjkError) as ei:
        fresh.upload_b(asym)
    assert "symmetric" in str(ei.value)
    fresh.upload_b(asym, sym_tol=1e-2)  # accepted when the caller allows it
    fresh.close()


def test_large_workload_properties(gpu_required):
    """nbf=1000, naux=3600, nbf5=300 at full size: size-independent identities."""
    import torch

    w = synth.workload(2)
    _say(f"case: {w.name} properties")
    ld = JKEngine.packed_ld(w.nbf, w.naux)
    btri = synth.device_b_packed(w.nbf, w.naux, w.seed_b, w.lam, ld)
    e = JKEngine(0)
    e.upload_b_packed(btri, w.nbf, w.naux)
    del btri
    C = torch.from_numpy(synth.make_C(w.nbf, w.seed_c)).cuda()
    H = torch.from_numpy(synth.make_H(w.nbf, w.seed_h)).cuda()
    r = e.build(C, w.nbf5, H=H)
    torch.cuda.synchronize()
    JM, KM = r["J_MO"], r["K_MO"]
    # J_MO = d d^T and K_MO = row sums of Y^2 come from different kernels; diag must agree
    assert torch.max(torch.abs(torch.diag(JM) - torch.diag(KM))) <= 1e-12 * torch.max(torch.abs(JM))
    assert torch.equal(JM, JM.T) and torch.equal(KM, KM.T)
    # J^q, K^q contracted with D_p = C_p C_p^T reproduce J_MO[q, p], K_MO[q, p] (integrals.py:337-341)
    Cq = C[:, : w.nbf5]
    for q in (0, 1, 150, 299):
        jq = Cq.T @ r["J"][q] @ Cq
        kq = Cq.T @ r["K"][q] @ Cq
        assert torch.max(torch.abs(torch.diag(jq) - JM[q])) <= 1e-10 * torch.max(torch.abs(JM))
        assert torch.max(torch.abs(torch.diag(kq) - KM[q])) <= 1e-10 * torch.max(torch.abs(KM))
        assert torch.equal(r["J"][q], r["J"][q].T) and torch.equal(r["K"][q], r["K"][q].T)
    hc = torch.einsum("mi,mi->i", Cq, H @ Cq)
    assert torch.max(torch.abs(hc - r["H_core"])) <= 1e-12 * torch.max(torch.abs(hc))
    e.close()
    del r
    torch.cuda.empty_cache()


def test_device_outputs_ordered_on_current_stream(eng):
    """A build with device outputs is ordered on torch's current stream: torch ops
    issued right after it (no explicit sync) read the complete outputs."""
    import torch

    print("case: stream ordering nbf=120 naux=400 nbf5=120", flush=True)
    nbf, naux, nbf5 = 120, 400, 120
    b = synth.make_b(nbf, naux, 31, 0.05)
    C = synth.make_C(nbf, 32)
    eng.upload_b(b)
    ref = eng.build(C, nbf5, outputs=("J_MO", "K_MO"))
    Cd = torch.from_numpy(C).cuda()
    want = float((2.0 * torch.from_numpy(ref["J_MO"]).cuda() - torch.from_numpy(ref["K_MO"]).cuda()).sum())
    for _ in range(3):
        out = eng.build(Cd, nbf5, outputs=("J", "K", "J_MO", "K_MO"))
        s = float((2.0 * out["J_MO"] - out["K_MO"]).sum())  # no synchronize before this
        assert s == want


def test_dropin_host_outputs_page_locked(gpu_required):
    """Host outputs of 64 MiB and more come back in page-locked memory (device-to-host at
    full PCIe rate) as plain numpy arrays, with the bits of a build into device outputs;
    the arrays own their memory for as long as they live."""
    import gc

    import torch

    nbf, naux, nbf5 = 220, 100, 180  # J, K 69.7 MB each
    print(f"case: drop-in host outputs pinned nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 41, 0.05)
    C = synth.make_C(nbf, 42)
    p = param(nbf, nbf5, nbfaux=naux, ndoc=1)
    J, K = computeJKj(C, [], b, p)
    for a in (J, K):
        assert isinstance(a, np.ndarray) and a.dtype == np.float64 and a.shape == (nbf5, nbf, nbf)
        assert a.flags.c_contiguous and a.flags.writeable
        assert torch.from_numpy(a).is_pinned()
    e = JKEngine(0)
    e.upload_b(b)
    d = e.build(torch.from_numpy(C).cuda(), nbf5, outputs=("J", "K"))
    torch.cuda.synchronize()
    assert np.array_equal(J, d["J"].cpu().numpy()) and np.array_equal(K, d["K"].cpu().numpy())
    assert rel(J, O.jkj_ri(C, b, nbf5)[0]) <= TOL
    # small outputs stay in ordinary numpy memory
    JM, KM, Hc = computeJKH_MO(C, synth.make_H(nbf, 43), [], b, p)
    assert not torch.from_numpy(JM).is_pinned()
    # an output that outlives the call keeps its contents after later builds reuse the pool
    Jc = J.copy()
    del K
    gc.collect()
    J2, K2 = computeJKj(C @ synth.rotation(nbf, 6001), [], b, p)
    assert np.array_equal(J, Jc) and not np.shares_memory(J, J2)
    e.close()


@pytest.mark.parametrize("nbf,naux,nbf5", [(60, 200, 40), (130, 400, 70), (97, 300, 96), (2, 3, 1), (40, 100, 40)])
def test_fp64_mo_only_builds(eng, nbf, naux, nbf5):
    """An FP64 build asked for J_MO, K_MO and H_core only forms the MO integrals by the
    triangular route (T = L C, W = C^T T) when nbf5 <= 2 nbf / 3, where it is the faster
    route: against the oracle at the FP64 bar, within 1e-12 of the full build (another
    summation order; past the threshold the full route is kept and the bits are the full
    build's), H_core bitwise, and the drop-in computeJKH_MO gives the same arrays."""
    import torch

    print(f"case: fp64 MO-only nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 51, 0.05)
    C = synth.make_C(nbf, 52)
    H = synth.make_H(nbf, 53)
    eng.upload_b(b)
    full = eng.build(C, nbf5, H=H)
    mo = eng.build(C, nbf5, H=H, outputs=("J_MO", "K_MO", "H_core"))
    JMr, KMr, Hr = O.jkh_mo_ri(C, H, b, nbf5)
    for k, r in (("J_MO", JMr), ("K_MO", KMr), ("H_core", Hr)):
        assert rel(mo[k], r) <= TOL, k
        assert np.linalg.norm(mo[k] - full[k]) <= 1e-12 * np.linalg.norm(full[k]), k
    assert np.array_equal(mo["H_core"], full["H_core"])
    if 3 * nbf5 > 2 * nbf:  # the full route
        for k in ("J_MO", "K_MO"):
            assert np.array_equal(mo[k], full[k]), k
    else:  # the route's reduction forms both diagonals in one order: d_p . d_p = sum Y_pp^2
        assert np.array_equal(np.diag(mo["J_MO"]), np.diag(mo["K_MO"]))
    # device inputs and outputs give the same bits as host ones
    dev = eng.build(torch.from_numpy(C).cuda(), nbf5, H=torch.from_numpy(H).cuda(),
                    outputs=("J_MO", "K_MO", "H_core"))
    torch.cuda.synchronize()
    for k in mo:
        assert np.array_equal(dev[k].cpu().numpy(), mo[k]), k
    p = param(nbf, nbf5, nbfaux=naux, ndoc=1)
    JM, KM, Hc = computeJKH_MO(C, H, [], b, p)
    assert np.array_equal(JM, mo["J_MO"]) and np.array_equal(KM, mo["K_MO"]) and np.array_equal(Hc, mo["H_core"])


def test_no_orbitals_gives_empty_outputs(gpu_required):
    """nbf5 = 0: zero-size outputs of the reference's shapes, as its slicing C[:, :0] gives
    (integrals.py:211, :351-360), for host and device C. (param itself refuses nbf5 = 0, as
    the reference's parameters do; a plain object with nbf5 and RI reaches the drop-ins.)"""
    import types

    import torch

    print("case: nbf5 = 0", flush=True)
    w = synth.workload(0)
    b, C, H, _ = synth.make_inputs(w)
    p = types.SimpleNamespace(nbf5=0, RI=True)
    J, K = computeJKj(C, [], b, p)
    JM, KM, Hc = computeJKH_MO(C, H, [], b, p)
    Jr, Kr = O.jkj_ri(C, b, 0)
    JMr, KMr, Hr = O.jkh_mo_ri(C, H, b, 0)
    for a, r in ((J, Jr), (K, Kr), (JM, JMr), (KM, KMr), (Hc, Hr)):
        assert isinstance(a, np.ndarray) and a.shape == r.shape and a.size == 0
    e = JKEngine(0)
    e.upload_b(b)
    d = e.build(torch.from_numpy(C).cuda(), 0, H=torch.from_numpy(H).cuda())
    assert d["J"].shape == (0, w.nbf, w.nbf) and d["J"].is_cuda and d["H_core"].shape == (0,)
    e.close()


@pytest.mark.parametrize("nbf,naux,nbf5", [(150, 400, 120), (1, 1, 1), (5, 7, 3), (40, 144, 40), (161, 96, 161),
                                          (352, 64, 337)])
@pytest.mark.parametrize("precision", ["fp64", "split", "split_refine"])
def test_jk_without_mo_outputs(eng, precision, nbf, naux, nbf5):
    """computeJKj's build (J and K, no J_MO or K_MO) runs the quarter transform for d alone
    (window row 0) and skips the K_MO reduction and J_MO: J and K keep the bits of the full
    build's (d has the same MFMA sequence either way); J alone and K alone as well. Shapes
    from one orbital to three M-tiles of orbitals and a two-tile window; split_refine keeps
    its FP64 d."""
    print(f"case: J, K without MO outputs {precision} nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 61, 0.05)
    C = synth.make_C(nbf, 62)
    H = synth.make_H(nbf, 63)
    eng.upload_b(b)
    full = eng.build(C, nbf5, H=H, precision=precision)
    jk = eng.build(C, nbf5, precision=precision, outputs=("J", "K"))
    assert np.array_equal(jk["J"], full["J"]) and np.array_equal(jk["K"], full["K"])
    j = eng.build(C, nbf5, precision=precision, outputs=("J",))
    k = eng.build(C, nbf5, precision=precision, outputs=("K",))
    assert np.array_equal(j["J"], full["J"]) and np.array_equal(k["K"], full["K"])
    if precision == "fp64" and nbf <= 200:
        Jr, Kr = O.jkj_ri(C, b, nbf5)
        assert rel(jk["J"], Jr) <= TOL and rel(jk["K"], Kr) <= TOL
