"""The committed water-sized fixture: the oracle reproduces it, and the seeded
generators reproduce its inputs."""
import hashlib
import json
import os

import numpy as np

from oracle import jk_oracle as O
from pynof_amd import synth

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def load():
    z = np.load(os.path.join(HERE, "water_fp64.npz"))
    with open(os.path.join(HERE, "water_fp64.json")) as f:
        meta = json.load(f)
    return {k: z[k] for k in z.files}, meta


def test_fixture_hashes():
    g, meta = load()
    for k, v in g.items():
        assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() == meta["sha256"][k], k


def test_generators_reproduce_inputs():
    g, meta = load()
    w = synth.workload(0)
    b, C, H, n = synth.make_inputs(w)
    assert np.array_equal(b, g["b"]) and np.array_equal(C, g["C"])
    assert np.array_equal(H, g["H"]) and np.array_equal(n, g["n"])


def test_oracle_reproduces_fixture():
    g, meta = load()
    nbf5 = meta["nbf5"]
    J, K = O.jkj_ri(g["C"], g["b"], nbf5)
    JM, KM, Hc = O.jkh_mo_ri(g["C"], g["H"], g["b"], nbf5)
    for name, a in (("J", J), ("K", K), ("J_MO", JM), ("K_MO", KM), ("H_core", Hc)):
        ref = g[name]
        assert np.max(np.abs(a - ref)) <= 1e-12 * np.max(np.abs(ref)), name
    e = O.energy(g["n"], JM, KM, Hc)
    assert abs(e - meta["E"]) < 1e-11 * abs(meta["E"])
