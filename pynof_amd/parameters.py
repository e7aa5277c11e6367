"""The subset of PyNOF's ``param`` object that the JK build reads.

The reference's ``param`` (src/pynof/parameters.py:5-112) is built from a psi4
wavefunction; the JK path reads only ``nbf``, ``nbf5``, ``nbfaux`` and the flags
``RI``, ``gpu`` and ``jit`` (parameters.py:16-17, 65, 102-104; integrals.py:44); the
Lagrangian and Fock calls add ``nalpha``, ``nbeta``, ``MSpin``, ``no1`` and ``noptorb``;
the occupation step adds ``ndoc``, ``nsoc``, ``ncwo``, ``nv``, ``ndns``, ``ipnof``,
``ista``, ``occ_method`` and ``HighSpin`` (parameters.py:46-48, 76-77, 91-94, 106, 110).
``param`` here is constructed from sizes. It carries the same attribute names,
plus ``precision`` ("fp64", "split", "split_refine") and ``device``.
"""
from __future__ import annotations


class param:
    def __init__(self, nbf: int, nbf5: int | None = None, nbfaux: int = 0, ndoc: int | None = None,
                 RI: bool = True, gpu: bool = True, jit: bool = False, precision: str = "fp64", device: int = 0,
                 no1: int = 0, noptorb: int | None = None, nsoc: int = 0, ncwo: int | None = None,
                 ipnof: int = 8, ista: int = 0, occ_method: str = "Softmax", HighSpin: bool = False):
        self.nbf = int(nbf)
        self.nbf5 = int(nbf if nbf5 is None else nbf5)
        if not 1 <= self.nbf5 <= self.nbf:
            raise ValueError("nbf5 must be in [1, nbf] (parameters.py:65 warns that nbf must exceed nbf5)")
        self.nbfaux = int(nbfaux)
        self.ndoc = ndoc
        self.no1 = int(no1)  # inactive doubly occupied orbitals (parameters.py:72)
        self.nsoc = int(nsoc)
        # nbeta = no1 + ndoc, nalpha = nbeta + nsoc (parameters.py:46-47)
        self.nbeta = None if ndoc is None else self.no1 + int(ndoc)
        self.nalpha = None if ndoc is None else self.nbeta + self.nsoc
        self.ndns = None if ndoc is None else int(ndoc) + self.nsoc  # parameters.py:48
        if ncwo is None and ndoc:
            ncwo = (self.nbf5 - self.no1 - self.nsoc - int(ndoc)) // int(ndoc)
        self.ncwo = ncwo  # weak orbitals per strong orbital (parameters.py:91)
        self.nv = None if ndoc is None or ncwo is None else int(ncwo) * int(ndoc)  # parameters.py:94
        self.MSpin = 0
        self.ipnof = int(ipnof)  # parameters.py:76
        self.ista = int(ista)  # parameters.py:77
        self.occ_method = occ_method  # parameters.py:110
        self.HighSpin = bool(HighSpin)  # parameters.py:106
        self.noptorb = int(self.nbf if noptorb is None else noptorb)  # parameters.py:68, :92
        self.RI = bool(RI)
        self.gpu = bool(gpu)
        self.jit = bool(jit)
        self.precision = precision
        self.device = int(device)
