"""Grouped builds (GPU only): bitwise equal to building each instance alone, in
fp64, split and split_refine forms, eagerly and from a replayed hipGraph."""
import numpy as np
import pytest

from pynof_amd import JKEngine, autotune, synth
from pynof_amd.batch import JKGroup

pytestmark = pytest.mark.gpu
SIZES = [(40, 144, 40), (97, 352, 60), (130, 472, 130), (64, 232, 50), (201, 728, 150)]
OUTS = ("J", "K", "J_MO", "K_MO", "H_core")


@pytest.fixture(scope="module")
def setup(gpu_required):
    import torch

    engs, Cs, Hs, n5s = [], [], [], []
    for i, (nbf, naux, n5) in enumerate(SIZES):
        e = JKEngine(0)
        e.upload_b(synth.make_b(nbf, naux, 900 + i, 0.05))
        engs.append(e)
        Cs.append(torch.from_numpy(synth.make_C(nbf, 950 + i)).cuda())
        Hs.append(torch.from_numpy(synth.make_H(nbf, 980 + i)).cuda())
        n5s.append(n5)
    yield engs, Cs, Hs, n5s
    for e in engs:
        e.close()


def alone(engs, Cs, Hs, n5s, prec):
    import torch

    res = []
    for e, c, h, n5 in zip(engs, Cs, Hs, n5s):
        r = e.build(c, n5, H=h, precision=prec)
        torch.cuda.synchronize()
        res.append({k: v.cpu().numpy() for k, v in r.items()})
    return res


@pytest.mark.parametrize("prec", ["fp64", "split", "split_refine"])
def test_group_equals_alone(setup, prec):
    import torch

    engs, Cs, Hs, n5s = setup
    print(f"case: group vs alone {prec} sizes {SIZES}", flush=True)
    ref = alone(engs, Cs, Hs, n5s, prec)
    g = JKGroup(engs, tuned={})
    out = g.build(Cs, n5s, Hs, precision=prec)
    torch.cuda.synchronize()
    for i, (o, r) in enumerate(zip(out, ref)):
        for k in OUTS:
            assert np.array_equal(o[k].cpu().numpy(), r[k]), (prec, i, k)
    g.close()


def test_group_graph_replays(setup):
    """Capture once, then replay with new C contents written in place; every
    replay matches an eager build of the same C, with eager work in between."""
    import torch

    engs, Cs, Hs, n5s = setup
    print("case: group graph replays split", flush=True)
    Cdev = [c.clone() for c in Cs]
    g = JKGroup(engs, tuned={})
    outs = g.alloc_outputs(n5s)
    for rep in range(3):
        for i, c in enumerate(Cdev):  # rotate C in place: same addresses, new values
            R = torch.from_numpy(synth.rotation(c.shape[0], 7000 + 10 * rep + i)).cuda()
            c.copy_(Cs[i] @ R)
        torch.cuda.synchronize()
        g.build(Cdev, n5s, Hs, precision="split", outs=outs, use_graph=True)
        torch.cuda.synchronize()
        got = [{k: v.cpu().numpy() for k, v in o.items()} for o in outs]
        ref = alone(engs, Cdev, Hs, n5s, "split")  # eager work between replays
        for i in range(len(engs)):
            for k in OUTS:
                assert np.array_equal(got[i][k], ref[i][k]), (rep, i, k)
    g.close()


def test_group_rejects_host_outputs(setup):
    import ctypes

    from pynof_amd import _lib

    engs, Cs, Hs, n5s = setup
    g = JKGroup(engs[:1], tuned={})
    J = np.empty((n5s[0], engs[0].nbf, engs[0].nbf))
    arr = (ctypes.c_void_p * 1)(J.ctypes.data)
    Carr = (ctypes.c_void_p * 1)(Cs[0].data_ptr())
    i64 = ctypes.c_int64 * 1
    rc = _lib.lib().pnjk_group_build(g._h, Carr, i64(Cs[0].shape[1]), i64(n5s[0]), None, None, 0, 1, arr, None,
                                     None, None, None, 0, None)
    assert rc == _lib.EINVAL
    g.close()


# Quarter-transform tiles 128 columns wide (fn = 4) as well as the small defaults: the
# reused plan's K_MO reduction must count 64-column partials for either width.
WIDE_QT = {"split:100:360:100": {"quarter_transform": [5, 4]}, "fp64:100:360:100": {"quarter_transform": [2, 4]}}


@pytest.mark.parametrize("tuned", [{}, WIDE_QT], ids=["default-tiles", "wide-qt"])
def test_group_new_C_tensors_reuse_plan(setup, tuned):
    """An optimiser hands a new C tensor each iteration: the group keeps its plan
    (eager) and re-captures its graph (use_graph), and every build matches eager
    single builds of the same C, bitwise."""
    import torch

    engs, Cs, Hs, n5s = setup
    print(f"case: group with new C tensors, eager then graph, split and fp64, tiles {tuned}", flush=True)
    g = JKGroup(engs, tuned=tuned)
    outs = g.alloc_outputs(n5s)
    keep = []  # earlier C tensors stay alive, so every iteration's C has new addresses
    for rep, (prec, graph) in enumerate((("split", False), ("split", False), ("fp64", False), ("fp64", False),
                                         ("split_refine", False), ("split", True), ("split", True))):
        Cnew = [Cs[i] @ torch.from_numpy(synth.rotation(c.shape[0], 7500 + 10 * rep + i)).cuda()
                for i, c in enumerate(Cs)]
        keep.append(Cnew)
        torch.cuda.synchronize()
        g.build(Cnew, n5s, Hs, precision=prec, outs=outs, use_graph=graph)
        torch.cuda.synchronize()
        got = [{k: v.cpu().numpy() for k, v in o.items()} for o in outs]
        ref = alone(engs, Cnew, Hs, n5s, prec)
        for i in range(len(engs)):
            for k in OUTS:
                assert np.array_equal(got[i][k], ref[i][k]), (rep, prec, graph, i, k)
    g.close()
    for e in engs:  # the group set these engines' tiles; back to the built-in choices
        autotune.apply(e, "split", None)
        autotune.apply(e, "fp64", None)


def test_group_replans_after_reupload_with_smaller_naux(gpu_required):
    """An engine of a group gets a new b with the same nbf and a smaller naux: its grow-only
    buffers (and so every pointer) stay, but the geometry differs, so the group must build a
    new plan; the grouped result equals a single build of the new b."""
    import torch

    print("case: group re-plan after re-upload with smaller naux", flush=True)
    sizes = [(64, 232, 64), (48, 176, 40)]
    engs = [JKEngine(0) for _ in sizes]
    for i, (nbf, naux, _) in enumerate(sizes):
        engs[i].upload_b(synth.make_b(nbf, naux, 1300 + i, 0.05))
    Cs = [torch.from_numpy(synth.make_C(n, 1310 + i)).cuda() for i, (n, _, _) in enumerate(sizes)]
    Hs = [torch.from_numpy(synth.make_H(n, 1320 + i)).cuda() for i, (n, _, _) in enumerate(sizes)]
    n5s = [s[2] for s in sizes]
    g = JKGroup(engs, tuned={})
    outs = g.alloc_outputs(n5s)
    for prec in ("fp64", "split"):
        g.build(Cs, n5s, Hs, precision=prec, outs=outs)
        torch.cuda.synchronize()
        engs[0].upload_b(synth.make_b(64, 104, 1330, 0.05))  # same nbf, naux 232 -> 104
        g.build(Cs, n5s, Hs, precision=prec, outs=outs)
        torch.cuda.synchronize()
        for i, e in enumerate(engs):
            r = e.build(Cs[i], n5s[i], H=Hs[i], precision=prec)
            torch.cuda.synchronize()
            for k in OUTS:
                assert torch.equal(outs[i][k], r[k]), (prec, i, k)
        engs[0].upload_b(synth.make_b(64, 232, 1300, 0.05))
    g.close()
    for e in engs:
        e.close()


SUBSETS = [("J", "K"), ("J_MO", "K_MO", "H_core"), ("K",), ("J", "J_MO"), ("K_MO",)]


@pytest.mark.parametrize("prec", ["fp64", "split", "split_refine"])
def test_group_output_subsets_equal_alone(setup, prec):
    """Grouped builds asking for part of the outputs (the J, K-only and MO-only routes
    change each instance's quarter transform and stages) equal single builds asking for
    the same outputs, bitwise; a group re-plans when the subset changes."""
    import torch

    engs, Cs, Hs, n5s = setup
    g = JKGroup(engs, tuned={})
    for sub in SUBSETS:
        print(f"case: group vs alone {prec} outputs {sub}", flush=True)
        hs = Hs if "H_core" in sub else None
        ref = []
        for e, c, h, n5 in zip(engs, Cs, Hs, n5s):
            r = e.build(c, n5, H=h if hs is not None else None, precision=prec, outputs=sub)
            torch.cuda.synchronize()
            ref.append({k: v.cpu().numpy() for k, v in r.items()})
        out = g.build(Cs, n5s, hs, precision=prec, outputs=sub)
        torch.cuda.synchronize()
        for i, (o, r) in enumerate(zip(out, ref)):
            assert set(o) == set(sub), (prec, sub, i, sorted(o))
            for k in sub:
                assert np.array_equal(o[k].cpu().numpy(), r[k]), (prec, sub, i, k)
    g.close()
