"""The subset of PyNOF's ``param`` object that the JK build reads.

The reference's ``param`` (src/pynof/parameters.py:5-112) is built from a psi4
wavefunction; the JK path reads only ``nbf``, ``nbf5``, ``nbfaux`` and the flags
``RI``, ``gpu`` and ``jit`` (parameters.py:16-17, 65, 102-104; integrals.py:44); the
Lagrangian and Fock calls add ``nalpha``, ``nbeta``, ``MSpin``, ``no1`` and ``noptorb``.
``param`` here is constructed from sizes. It carries the same attribute names,
plus ``precision`` ("fp64", "split", "split_refine") and ``device``.
"""
from __future__ import annotations


class param:
    def __init__(self, nbf: int, nbf5: int | None = None, nbfaux: int = 0, ndoc: int | None = None,
                 RI: bool = True, gpu: bool = True, jit: bool = False, precision: str = "fp64", device: int = 0,
                 no1: int = 0, noptorb: int | None = None):
        self.nbf = int(nbf)
        self.nbf5 = int(nbf if nbf5 is None else nbf5)
        if not 1 <= self.nbf5 <= self.nbf:
            raise ValueError("nbf5 must be in [1, nbf] (parameters.py:65 warns that nbf must exceed nbf5)")
        self.nbfaux = int(nbfaux)
        self.ndoc = ndoc
        self.nalpha = self.nbeta = ndoc
        self.MSpin = 0
        self.no1 = int(no1)  # inactive doubly occupied orbitals (parameters.py:72)
        self.noptorb = int(self.nbf if noptorb is None else noptorb)  # parameters.py:68, :92
        self.RI = bool(RI)
        self.gpu = bool(gpu)
        self.jit = bool(jit)
        self.precision = precision
        self.device = int(device)
