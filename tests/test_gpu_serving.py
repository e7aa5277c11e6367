"""Serving loop on the GPU: every request's energies and final MO integrals are
the same alone and in the mix (bitwise: grouped builds equal single builds and
each request's rotations are seeded per request), and a 1e-9 Eh request's final
energy equals the FP64 form's energy at its final orbitals."""
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
    for k in ("J_MO", "K_MO", "H_core"):
        assert np.array_equal(o[k], r.final_outs[k]), k
    e1, e2 = synth.energy_terms(n, o["J_MO"], o["K_MO"], o["H_core"])
    assert abs((e1 + e2) - r.energies[-1]) <= 1e-10 * abs(e1 + e2)
    eng.close()
