"""pnjk_jk_build_noalloc (workspace in the library's static scratch array, no
allocation) against pnjk_jk_build on the same engine and inputs, through the C ABI:
bitwise equal outputs in every form, on two calls in a row (the second on a smaller
instance, so the shared scratch holds the first call's data), and a clear error for a
build whose workspace exceeds the array."""
import ctypes

import numpy as np
import pytest

from pynof_amd import JKEngine, _lib, synth

pytestmark = pytest.mark.gpu
FORMS = ("fp64", "split", "split_refine")


def _outs(torch, nbf, n5):
    return {"J": torch.empty((n5, nbf, nbf), dtype=torch.float64, device="cuda"),
            "K": torch.empty((n5, nbf, nbf), dtype=torch.float64, device="cuda"),
            "J_MO": torch.empty((n5, n5), dtype=torch.float64, device="cuda"),
            "K_MO": torch.empty((n5, n5), dtype=torch.float64, device="cuda"),
            "H_core": torch.empty((n5,), dtype=torch.float64, device="cuda")}


def _call(fn, eng, C, n5, H, prec, o):
    torch = pytest.importorskip("torch")
    s = torch.cuda.current_stream().cuda_stream
    return fn(eng._h, C.data_ptr(), C.shape[1], n5, H.data_ptr(), H.shape[1], _lib.PRECISIONS[prec], _lib.OUT_ALL,
              o["J"].data_ptr(), o["K"].data_ptr(), o["J_MO"].data_ptr(), o["K_MO"].data_ptr(),
              o["H_core"].data_ptr(), ctypes.c_void_p(s))


def test_noalloc_equals_build_two_calls(gpu_required):
    import torch

    L = _lib.lib()
    cases = [(88, 320, 80, 21), (40, 144, 30, 22)]  # second call: a smaller instance
    engs = []
    for nbf, naux, n5, seed in cases:
        e = JKEngine(0)
        e.upload_b(synth.make_b(nbf, naux, seed, 0.05))
        engs.append(e)
    for prec in FORMS:
        for (nbf, naux, n5, seed), e in zip(cases, engs):
            print(f"case: noalloc {prec} nbf={nbf} naux={naux} nbf5={n5}", flush=True)
            C = torch.from_numpy(synth.make_C(nbf, seed + 100)).cuda()
            H = torch.from_numpy(synth.make_H(nbf, seed + 200)).cuda()
            ref, got = _outs(torch, nbf, n5), _outs(torch, nbf, n5)
            _lib.check(_call(L.pnjk_jk_build, e, C, n5, H, prec, ref))  # also makes the split planes
            for v in got.values():
                v.fill_(float("nan"))
            _lib.check(_call(L.pnjk_jk_build_noalloc, e, C, n5, H, prec, got))
            torch.cuda.synchronize()
            for k in ref:
                assert torch.equal(ref[k], got[k]), (prec, nbf, k)
    for e in engs:
        e.close()


def test_noalloc_rejects_oversized_and_host_outputs(gpu_required):
    import torch

    L = _lib.lib()
    nbf, naux = 200, 720
    e = JKEngine(0)
    e.upload_b(synth.make_b(nbf, naux, 31, 0.05))
    C = torch.from_numpy(synth.make_C(nbf, 32)).cuda()
    H = torch.from_numpy(synth.make_H(nbf, 33)).cuda()
    o = _outs(torch, nbf, nbf)
    rc = _call(L.pnjk_jk_build_noalloc, e, C, nbf, H, "fp64", o)
    assert rc == _lib.EINVAL and b"bytes of workspace" in L.pnjk_last_error()
    Jh = np.empty((nbf, nbf, nbf))
    s = torch.cuda.current_stream().cuda_stream
    rc = L.pnjk_jk_build_noalloc(e._h, C.data_ptr(), nbf, 8, None, 0, _lib.PRECISIONS["fp64"], _lib.OUT_J,
                                 Jh.ctypes.data, None, None, None, None, ctypes.c_void_p(s))
    assert rc == _lib.EINVAL and b"device memory" in L.pnjk_last_error()
    e.close()


def test_noalloc_fp64_mo_only_takes_the_triangular_route(gpu_required):
    """An FP64 build without J and K at nbf5 <= 2 nbf / 3 takes the triangular MO route through
    both entry points: the no-allocation build refuses it until the halved diagonal rows of
    b exist (the first such pnjk_jk_build makes them), then equals pnjk_jk_build bitwise, on
    two calls in a row (the second on a smaller instance). The split form's triangular route
    likewise, once the split planes exist."""
    import torch

    L = _lib.lib()
    mo = _lib.OUT_JMO | _lib.OUT_KMO | _lib.OUT_HCORE
    s = torch.cuda.current_stream().cuda_stream
    for nbf, naux, n5, seed in [(96, 352, 36, 41), (48, 176, 18, 42)]:  # nbf5 <= 2 nbf / 5: both forms' route
        print(f"case: noalloc fp64 MO-only nbf={nbf} naux={naux} nbf5={n5}", flush=True)
        e = JKEngine(0)
        e.upload_b(synth.make_b(nbf, naux, seed, 0.05))
        C = torch.from_numpy(synth.make_C(nbf, seed + 100)).cuda()
        H = torch.from_numpy(synth.make_H(nbf, seed + 200)).cuda()
        ref, got = _outs(torch, nbf, n5), _outs(torch, nbf, n5)

        def call(fn, o):
            return fn(e._h, C.data_ptr(), C.shape[1], n5, H.data_ptr(), H.shape[1], _lib.FP64, mo, None, None,
                      o["J_MO"].data_ptr(), o["K_MO"].data_ptr(), o["H_core"].data_ptr(), ctypes.c_void_p(s))

        rc = call(L.pnjk_jk_build_noalloc, got)
        assert rc == _lib.ESTATE and b"halved diagonal" in L.pnjk_last_error()
        _lib.check(call(L.pnjk_jk_build, ref))
        for v in got.values():
            v.fill_(float("nan"))
        _lib.check(call(L.pnjk_jk_build_noalloc, got))
        torch.cuda.synchronize()
        for k in ("J_MO", "K_MO", "H_core"):
            assert torch.equal(ref[k], got[k]), (nbf, k)

        def call16(fn, o):
            return fn(e._h, C.data_ptr(), C.shape[1], n5, H.data_ptr(), H.shape[1], _lib.SPLIT, mo, None, None,
                      o["J_MO"].data_ptr(), o["K_MO"].data_ptr(), o["H_core"].data_ptr(), ctypes.c_void_p(s))

        print(f"case: noalloc split MO-only nbf={nbf} naux={naux} nbf5={n5}", flush=True)
        _lib.check(call16(L.pnjk_jk_build, ref))  # also makes the split planes
        for v in got.values():
            v.fill_(float("nan"))
        _lib.check(call16(L.pnjk_jk_build_noalloc, got))
        torch.cuda.synchronize()
        for k in ("J_MO", "K_MO", "H_core"):
            assert torch.equal(ref[k], got[k]), ("split", nbf, k)
        e.close()
