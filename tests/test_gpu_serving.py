"""Serving loop on the GPU: every request's energies and final MO integrals are
the same alone and in the mix (bitwise: grouped builds equal single builds and
each request's rotations are seeded per request), a 1e-9 Eh request's final
energy is within 1e-9 Eh of the FP64 form's energy at its final orbitals, and on the Poisson
request mix every request's final energy is within its own tolerance (Eh) of
the FP64 form's."""
import copy

import numpy as np
import pytest

from pynof_amd import JKEngine, synth
from pynof_amd.serving import DeviceExecutor, Request, Server

pytestmark = pytest.mark.gpu


def reqs():
    sizes = [(40, 0.0, 1e-9), (96, 0.0, 1e-6), (64, 0.02, 1e-9), (130, 0.05, 1e-9)]
    return [Request(rid=i, nbf=n, naux=synth.ceil8(3.6 * n), nbf5=n, ndoc=synth.default_ndoc(n), tol=tol,
                    n_builds=5, seed=8000 + 13 * i, arrival=t) for i, (n, t, tol) in enumerate(sizes)]


def serve(rs):
    srv = Server(DeviceExecutor(0, keep_final_C=True), hbm_budget=200e9)
    for r in rs:
        srv.submit(r)
    srv.run(timeout=300)
    assert not srv.rejected
    return {r.rid: r for r in srv.finished}, srv


def test_mix_equals_alone(gpu_required):
    print("case: serving mix vs alone, 4 requests", flush=True)
    mix, srv = serve(reqs())
    assert max(srv.step_sizes) >= 2  # requests did share steps
    for r0 in reqs():
        r = copy.deepcopy(r0)
        r.arrival = 0.0
        alone, _ = serve([r])
        a, m = alone[r0.rid], mix[r0.rid]
        assert a.energies == m.energies, r0.rid
        for k in ("J_MO", "K_MO", "H_core", "J", "K"):
            assert np.array_equal(a.final_outs[k], m.final_outs[k]), (r0.rid, k)


def test_pooled_engine_reuse_equals_fresh(gpu_required):
    """A request served on a pooled engine (left by a larger request, whose b, split planes
    and refine workspace it held) gives the bits of the same request on a fresh engine."""
    print("case: serving pooled engine reuse vs fresh", flush=True)
    big = Request(rid=0, nbf=104, naux=synth.ceil8(3.6 * 104), nbf5=104, ndoc=synth.default_ndoc(104), tol=1e-9,
                  n_builds=3, seed=8100)
    small = Request(rid=1, nbf=96, naux=synth.ceil8(3.6 * 96), nbf5=96, ndoc=synth.default_ndoc(96), tol=1e-9,
                    n_builds=3, seed=8200)
    ex = DeviceExecutor(0, keep_final_C=True)
    srv = Server(ex, hbm_budget=200e9)
    srv.submit(big)
    srv.run(timeout=300)
    assert len(ex._idle) == 1  # the larger request's engine is idle in the pool
    pooled = ex._idle[0][1]
    srv = Server(ex, hbm_budget=200e9)
    srv.submit(small)
    srv.run(timeout=300)
    assert ex._idle[0][1] is pooled  # the small request ran on it and returned it
    fresh, _ = serve([copy.deepcopy(Request(rid=1, nbf=96, naux=synth.ceil8(3.6 * 96), nbf5=96,
                                            ndoc=synth.default_ndoc(96), tol=1e-9, n_builds=3, seed=8200))])
    f = fresh[1]
    assert small.energies == f.energies
    for k in ("J_MO", "K_MO", "H_core", "J", "K"):
        assert np.array_equal(small.final_outs[k], f.final_outs[k]), k
    ex.close()


def test_tight_request_final_energy_is_fp64(gpu_required):
    import torch

    print("case: serving 1e-9 request final energy vs fp64", flush=True)
    mix, _ = serve(reqs()[:1])
    r = mix[0]
    eng = JKEngine(0)
    btri = synth.device_b_packed(r.nbf, r.naux, r.seed, 0.05, JKEngine.packed_ld(r.nbf, r.naux))
    eng.upload_b_packed(btri, r.nbf, r.naux)
    H = synth.make_H(r.nbf, r.seed + 1)
    n = synth.make_n(r.nbf5, r.ndoc, r.seed + 2)
    o = eng.build(r.final_C, r.nbf5, H=H, precision="fp64", outputs=("J_MO", "K_MO", "H_core"))
    torch.cuda.synchronize()
    # the refined build's MO integrals come from the triangular FP64 route: within 1e-12 of
    # the FP64 form's (H_core bitwise), and the request's energy within 1e-9 Eh
    for k in ("J_MO", "K_MO"):
        assert np.linalg.norm(o[k] - r.final_outs[k]) <= 1e-12 * np.linalg.norm(o[k]), k
    assert np.array_equal(o["H_core"], r.final_outs["H_core"])
    e1, e2 = synth.energy_terms(n, o["J_MO"], o["K_MO"], o["H_core"])
    assert abs((e1 + e2) - r.energies[-1]) <= 1e-9
    eng.close()


def test_serving_with_occupation_step(gpu_required):
    """Each build followed by a device occupation step (GNOF, softmax): the final energy
    equals the oracle's calcocce at the request's final gamma and final MO integrals,
    and the mix equals a request alone."""
    from oracle import pnof_oracle as PO

    print("case: serving with the occupation step", flush=True)
    rs = reqs()[:3]
    srv = Server(DeviceExecutor(0, keep_final_C=True, occupation=True), hbm_budget=200e9)
    for r in rs:
        srv.submit(r)
    srv.run(timeout=300)
    assert not srv.rejected
    mix = {r.rid: r for r in srv.finished}
    for r in mix.values():
        ncwo = (r.nbf5 - r.ndoc) // r.ndoc
        m = r.ndoc * (1 + ncwo)
        p = PO.Pnof(0, r.ndoc, 0, ncwo, ipnof=8, occ_method="Softmax")
        o = r.final_outs
        E = PO.calcocce(r.final_gamma, o["J_MO"][:m, :m], o["K_MO"][:m, :m], o["H_core"][:m], p)
        assert abs(E - r.energies[-1]) <= 1e-12 * max(1.0, abs(E)), r.rid
        assert np.all(r.final_gamma != -2.0)  # the parameters moved on the device
    r0 = reqs()[1]  # a fresh request object: a served one is not resubmitted
    r0.arrival = 0.0
    srv1 = Server(DeviceExecutor(0, keep_final_C=True, occupation=True), hbm_budget=200e9)
    srv1.submit(r0)
    srv1.run(timeout=300)
    assert srv1.finished[0].energies == mix[r0.rid].energies


def test_poisson_mix_tolerances_and_alone(gpu_required):
    """The Poisson request mix (seed 5000: lognormal nbf, median 150; 20 builds with C
    rotated by exp(kappa) between builds, minimization.py:200-253, utils.py:550-565),
    capped at its first 24 requests, arriving at 200 req/s so that they overlap. Every request's
    final energy is within its own tolerance (Eh, absolute) of the FP64 form's energy at
    its final orbitals, and every request's energies are the same in the mix as alone, with
    lockstep steps and with size-class lanes."""
    import torch

    from pynof_amd.serving import poisson_requests, split_error_bound

    print("case: serving poisson seed 5000, 24 requests at 200 req/s", flush=True)
    keep = ("J_MO", "K_MO", "H_core")
    rs = poisson_requests(24, rate=200.0, seed=5000)
    assert {r.tol for r in rs} == {1e-6, 1e-9}
    srv = Server(DeviceExecutor(0, keep_final_C=True, keep_outs=keep), hbm_budget=200e9)
    for r in rs:
        srv.submit(r)
    srv.run(timeout=600)
    assert not srv.rejected and len(srv.finished) == 24
    print(f"steps {srv.steps}, largest step {max(srv.step_sizes)} requests", flush=True)
    assert max(srv.step_sizes) >= 2  # requests of different sizes shared steps
    mix = {r.rid: r for r in srv.finished}
    for r in sorted(mix.values(), key=lambda x: x.rid):
        assert len(r.energies) == r.n_builds
        refined = r.form_for_build(r.n_builds - 1) == "split_refine"
        assert refined == (r.tol < split_error_bound(r.nbf))
        eng = JKEngine(0)
        btri = synth.device_b_packed(r.nbf, r.naux, r.seed, 0.05, JKEngine.packed_ld(r.nbf, r.naux))
        eng.upload_b_packed(btri, r.nbf, r.naux)
        del btri
        H = synth.make_H(r.nbf, r.seed + 1)
        n = synth.make_n(r.nbf5, r.ndoc, r.seed + 2)
        o = eng.build(r.final_C, r.nbf5, H=H, precision="fp64", outputs=keep)
        torch.cuda.synchronize()
        e64 = sum(synth.energy_terms(n, o["J_MO"], o["K_MO"], o["H_core"]))
        err = abs(r.energies[-1] - e64)
        print(f"rid {r.rid} nbf {r.nbf} tol {r.tol:g} refined {refined} final |E - E_fp64| = {err:.3e} Eh", flush=True)
        assert err <= r.tol, (r.rid, r.nbf, r.tol, err)
        eng.close()
    # the same mix with size-class lanes (each class steps on its own stream): same results
    srv2 = Server(DeviceExecutor(0, keep_final_C=True, keep_outs=keep), hbm_budget=200e9, lanes=(128, 256))
    for r in poisson_requests(24, rate=200.0, seed=5000):
        srv2.submit(r)
    srv2.run(timeout=600)
    assert not srv2.rejected and len(srv2.finished) == 24
    assert {r.lane for r in srv2.finished} == {0, 1, 2}
    for r in srv2.finished:
        assert r.energies == mix[r.rid].energies, r.rid
        for k in keep:
            assert np.array_equal(r.final_outs[k], mix[r.rid].final_outs[k]), (r.rid, k)
    for r0 in poisson_requests(24, rate=200.0, seed=5000):
        r0.arrival = 0.0
        s1 = Server(DeviceExecutor(0, keep_final_C=True, keep_outs=keep), hbm_budget=200e9)
        s1.submit(r0)
        s1.run(timeout=300)
        a = s1.finished[0]
        assert a.energies == mix[r0.rid].energies, r0.rid
        for k in keep:
            assert np.array_equal(a.final_outs[k], mix[r0.rid].final_outs[k]), (r0.rid, k)
    torch.cuda.empty_cache()
