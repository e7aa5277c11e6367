"""The 64-instance batch (nbf ~ U[40, 400] from seed 4000, naux = ceil8(3.6 nbf),
nbf5 = nbf; integrals.py:172 and :307 build one system per call) through one
grouped build, eagerly and replayed from a captured hipGraph, in every form:

* every instance's outputs are bitwise those of building it alone;
* instances with nbf <= 130 are checked against the oracle (b read back from the
  device in the reference layout): FP64 within 1e-10 normwise relative, split within
  3e-5, split_refine's J_MO, K_MO, H_core within 1e-10.
"""
import numpy as np
import pytest

from oracle import jk_oracle as O
from pynof_amd import JKEngine, synth
from pynof_amd.batch import JKGroup

pytestmark = pytest.mark.gpu
OUTS = ("J", "K", "J_MO", "K_MO", "H_core")
ORACLE_MAX_NBF = 130


def rel(a, r):
    return float(np.max(np.abs(np.asarray(a) - r)) / max(1e-300, np.max(np.abs(r))))


@pytest.fixture(scope="module")
def batch(gpu_required):
    import torch

    ws = synth.batch_workloads(64)
    engs, Cs, Hs = [], [], []
    for w in ws:
        e = JKEngine(0)
        bt = synth.device_b_packed(w.nbf, w.naux, w.seed_b, w.lam, JKEngine.packed_ld(w.nbf, w.naux))
        e.upload_b_packed(bt, w.nbf, w.naux)
        del bt
        engs.append(e)
        Cs.append(torch.from_numpy(synth.make_C(w.nbf, w.seed_c)).cuda())
        Hs.append(torch.from_numpy(synth.make_H(w.nbf, w.seed_h)).cuda())
    torch.cuda.empty_cache()
    yield ws, engs, Cs, Hs
    for e in engs:
        e.close()
    torch.cuda.empty_cache()


def check_against_single(ws, engs, Cs, Hs, outs, prec, tag):
    import torch

    for i, (w, e) in enumerate(zip(ws, engs)):
        r = e.build(Cs[i], w.nbf5, H=Hs[i], precision=prec)
        torch.cuda.synchronize()
        for k in OUTS:
            assert torch.equal(outs[i][k], r[k]), (tag, prec, i, w.nbf, k)
        del r


@pytest.mark.parametrize("prec", ["fp64", "split", "split_refine"])
def test_batch64_grouped_and_graph_equal_single(batch, prec):
    import torch

    ws, engs, Cs, Hs = batch
    n5s = [w.nbf5 for w in ws]
    print(f"case: batch64 seed 4000 grouped eager + graph, {prec}", flush=True)
    g = JKGroup(engs)
    outs = g.alloc_outputs(n5s)
    g.build(Cs, n5s, Hs, precision=prec, outs=outs)
    torch.cuda.synchronize()
    check_against_single(ws, engs, Cs, Hs, outs, prec, "eager")
    for rep in range(2):  # capture, then a replay; outputs poisoned before each
        for o in outs:
            for v in o.values():
                v.fill_(float("nan"))
        g.build(Cs, n5s, Hs, precision=prec, outs=outs, use_graph=True)
        torch.cuda.synchronize()
        check_against_single(ws, engs, Cs, Hs, outs, prec, f"graph{rep}")
    small = [i for i, w in enumerate(ws) if w.nbf <= ORACLE_MAX_NBF]
    assert len(small) == 15
    for i in small:
        w = ws[i]
        b = engs[i].download_b()
        ref = dict(zip(OUTS, O.jk_build(Cs[i].cpu().numpy(), Hs[i].cpu().numpy(), b, w.nbf5)))
        got = {k: v.cpu().numpy() for k, v in outs[i].items()}
        for k in OUTS:
            err = rel(got[k], ref[k])
            if prec == "fp64" or (prec == "split_refine" and k in ("J_MO", "K_MO")) or k == "H_core":
                assert err <= 1e-10, (prec, i, w.nbf, k, err)
            else:
                assert err <= 3e-5, (prec, i, w.nbf, k, err)
    g.close()
    del outs
    torch.cuda.empty_cache()
