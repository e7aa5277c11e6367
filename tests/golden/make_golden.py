"""Writes the golden fixture of the water-sized workload (nbf=24, naux=116, nbf5=24).

Inputs come from pynof_amd.synth's seeded generators; outputs from the FP64
oracle with the reference's call sequence (oracle.jk_oracle.jkj_ri /
jkh_mo_ri, which follow integrals.py:208-218 and :348-362). The reference
itself cannot run here (psi4 and numba are absent), so the fixture is the
oracle's output, pinned by tests/test_oracle.py.

Run from the repository root:  python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import jk_oracle as O  # noqa: E402
from pynof_amd import synth  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))


def main():
    w = synth.workload(0)
    b, C, H, n = synth.make_inputs(w)
    J, K = O.jkj_ri(C, b, w.nbf5)
    JM, KM, Hc = O.jkh_mo_ri(C, H, b, w.nbf5)
    e1, e2 = synth.energy_terms(n, JM, KM, Hc)
    arrays = dict(b=b, C=C, H=H, n=n, J=J, K=K, J_MO=JM, K_MO=KM, H_core=Hc)
    np.savez_compressed(os.path.join(HERE, "water_fp64.npz"), **arrays)
    meta = {
        "workload": w.name,
        "nbf": w.nbf, "naux": w.naux, "nbf5": w.nbf5, "ndoc": w.ndoc, "lam": w.lam,
        "seeds": {"b": w.seed_b, "C": w.seed_c, "n": w.seed_n, "H": w.seed_h},
        "E1": e1, "E2": e2, "E": e1 + e2,
        "sha256": {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in arrays.items()},
        "numpy": np.__version__,
    }
    with open(os.path.join(HERE, "water_fp64.json"), "w") as f:
        json.dump(meta, f, indent=1)
    print(json.dumps({k: meta[k] for k in ("E1", "E2", "E")}))


if __name__ == "__main__":
    main()
