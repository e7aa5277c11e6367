"""Host-side engine: one device-resident b tensor and the builds against it.

``JKEngine`` wraps one ``pnjk_engine`` of the C ABI. b is uploaded once (the
reference's ``cp.array(b_mnl)``, integrals.py:51-53) and stays on the device in
the packed lower-triangle layout; each ``build`` call runs the fused JK build
(half-transform once, integrals.py:211 and :353) for one set of orbitals C.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _lib


def _is_torch(x) -> bool:
    return type(x).__module__.startswith("torch")


def _torch():
    import torch

    return torch


# Host outputs from this size up are page-locked: device-to-host copies into pageable
# memory run at 9-17 GB/s on the MI355X box, into pinned memory at about 50 GB/s (J and K
# at nbf = 1000, nbf5 = 300: 4.8 GB in 95 ms instead of 280-520 ms). Up to the cap only, so
# that a download of a whole b does not pin tens of GB.
PINNED_MIN_BYTES = 64 << 20
PINNED_MAX_BYTES = 16 << 30


def _host_empty(shape):
    """A new host FP64 array for a device-to-host result: a numpy view of a page-locked
    torch tensor (it lives as long as the array; torch's host allocator reuses its block
    afterwards) when the size is in [PINNED_MIN_BYTES, PINNED_MAX_BYTES] and the GPU is
    usable, else plain numpy memory."""
    nbytes = 8 * int(np.prod(shape))
    if PINNED_MIN_BYTES <= nbytes <= PINNED_MAX_BYTES:
        torch = _torch()
        if torch.cuda.is_available():
            return torch.empty(tuple(shape), dtype=torch.float64, pin_memory=True).numpy()
    return np.empty(shape, dtype=np.float64)


def _ptr(x) -> int:
    if _is_torch(x):
        return x.data_ptr()
    return x.ctypes.data


def _as_input(x, name: str):
    """C-contiguous FP64 view of a host (numpy) or device (torch) array."""
    if _is_torch(x):
        if x.dtype != _torch().float64:
            raise TypeError(f"{name} must be float64, got {x.dtype}")
        return x.contiguous()
    a = np.asarray(x)
    if a.dtype != np.float64:
        raise TypeError(f"{name} must be float64, got {a.dtype}")
    return np.ascontiguousarray(a)


OUTPUT_BITS = {"J": _lib.OUT_J, "K": _lib.OUT_K, "J_MO": _lib.OUT_JMO, "K_MO": _lib.OUT_KMO,
               "H_core": _lib.OUT_HCORE}


class JKEngine:
    """A device-resident b tensor and the JK builds that read it."""

    def __init__(self, device: int = 0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.pnjk_create(int(device), ctypes.byref(h)))
        self._h = h
        self.device = int(device)
        self.nbf = 0
        self.naux = 0
        self.asym = None

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().pnjk_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def footprint(self) -> int:
        return int(_lib.lib().pnjk_footprint(self._h))

    @staticmethod
    def packed_ld(nbf: int, naux: int) -> int:
        return int(_lib.lib().pnjk_packed_ld(int(nbf), int(naux)))

    def upload_b(self, b_mnl, sym_tol: float = 1e-10) -> int:
        """Uploads b (nbf, nbf, naux) from host or device memory; returns the footprint."""
        b = _as_input(b_mnl, "b_mnl")
        if b.ndim != 3 or b.shape[0] != b.shape[1]:
            raise ValueError(f"b_mnl must have shape (nbf, nbf, naux); got {tuple(b.shape)}")
        nbf, _, naux = (int(s) for s in b.shape)
        fp = ctypes.c_uint64()
        asym = ctypes.c_double()
        if _is_torch(b):
            _torch().cuda.current_stream(b.device).synchronize()
        self.nbf = 0  # the library holds no b until the call succeeds
        _lib.check(_lib.lib().pnjk_upload_b(self._h, _ptr(b), nbf, naux, float(sym_tol), ctypes.byref(fp),
                                            ctypes.byref(asym)))
        self.nbf, self.naux, self.asym = nbf, naux, asym.value
        return int(fp.value)

    def upload_b_packed(self, btri, nbf: int, naux: int) -> int:
        """Adopts (copies) a device tensor already in the packed layout."""
        ld = self.packed_ld(nbf, naux)
        if tuple(btri.shape) != (nbf * (nbf + 1) // 2, ld):
            raise ValueError(f"packed b must have shape {(nbf * (nbf + 1) // 2, ld)}; got {tuple(btri.shape)}")
        _torch().cuda.current_stream(btri.device).synchronize()
        fp = ctypes.c_uint64()
        self.nbf = 0  # the library holds no b until the call succeeds
        _lib.check(_lib.lib().pnjk_upload_b_packed(self._h, btri.data_ptr(), nbf, naux, ld, ctypes.byref(fp)))
        self.nbf, self.naux = nbf, naux
        return int(fp.value)

    def upload_b_3c(self, mnk, metric, cutoff: float = 1e-14, sym_tol: float = 1e-10) -> int:
        """Forms b on the device from the 3-centre integrals mnk[mu, nu, P] = (mu nu|P)
        (nbf, nbf, naux) and the metric (P|Q) (naux, naux) (integrals.py:35-49): b = mnk
        (P|Q)^-1/2 with the inverse square root cut as psi4's Matrix::power(-0.5, cutoff).
        Host (numpy) or device (torch) arrays. Returns the resident footprint in bytes."""
        m3 = _as_input(mnk, "mnk")
        mt = _as_input(metric, "metric")
        if m3.ndim != 3 or m3.shape[0] != m3.shape[1]:
            raise ValueError(f"mnk must have shape (nbf, nbf, naux); got {tuple(m3.shape)}")
        nbf, _, naux = (int(x) for x in m3.shape)
        if tuple(mt.shape) != (naux, naux):
            raise ValueError(f"metric must have shape ({naux}, {naux}); got {tuple(mt.shape)}")
        for x in (m3, mt):
            if _is_torch(x):
                _torch().cuda.current_stream(x.device).synchronize()
        fp = ctypes.c_uint64()
        asym = ctypes.c_double()
        self.nbf = 0  # the library holds no b until the call succeeds
        _lib.check(_lib.lib().pnjk_upload_b_3c(self._h, _ptr(m3), nbf, naux, _ptr(mt), float(cutoff),
                                               float(sym_tol), ctypes.byref(fp), ctypes.byref(asym)))
        self.nbf, self.naux, self.asym = nbf, naux, asym.value
        return int(fp.value)

    def download_b(self, out=None):
        """The resident b in the reference layout (nbf, nbf, naux): into `out` (numpy, a
        numpy memmap or a device tensor) or a new host array."""
        if self.nbf == 0:
            raise _lib.PnjkError(_lib.ESTATE, "no b uploaded")
        shape = (self.nbf, self.nbf, self.naux)
        if out is None:
            out = _host_empty(shape)
        if tuple(out.shape) != shape:
            raise ValueError(f"out must have shape {shape}; got {tuple(out.shape)}")
        if _is_torch(out):
            if out.dtype != _torch().float64 or not out.is_contiguous():
                raise ValueError("out must be a contiguous float64 tensor")
            _torch().cuda.current_stream(out.device).synchronize()
        elif out.dtype != np.float64 or not out.flags.c_contiguous:
            raise ValueError("out must be a C-contiguous float64 array")
        _lib.check(_lib.lib().pnjk_download_b(self._h, _ptr(out)))
        return out

    def build(self, C, nbf5: int, H=None, precision: str = "fp64",
              outputs=("J", "K", "J_MO", "K_MO", "H_core"), out=None, stream=None):
        """One fused JK build.

        C: (nbf, >= nbf5) orbitals, host numpy or device torch, FP64.
        H: (nbf, nbf) core Hamiltonian, needed for H_core.
        out: optional dict name -> preallocated array (numpy or device torch) to
        write into; missing outputs are allocated as numpy (synchronous call) or,
        when C is a device tensor, as device tensors (asynchronous on `stream`).
        Returns a dict name -> array.
        """
        if self.nbf == 0:
            raise _lib.PnjkError(_lib.ESTATE, "no b uploaded")
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        nbf5 = int(nbf5)
        Ci = _as_input(C, "C")
        if Ci.ndim != 2 or Ci.shape[0] != self.nbf or Ci.shape[1] < nbf5:
            raise ValueError(f"C must have shape ({self.nbf}, >= {nbf5}); got {tuple(Ci.shape)}")
        want = list(outputs)
        mask = 0
        for name in want:
            mask |= OUTPUT_BITS[name]
        Hi = None
        if "H_core" in want:
            if H is None:
                raise ValueError("H_core requested but H is None")
            Hi = _as_input(H, "H")
            if tuple(Hi.shape) != (self.nbf, self.nbf):
                raise ValueError(f"H must have shape ({self.nbf}, {self.nbf}); got {tuple(Hi.shape)}")
        on_device = _is_torch(Ci) and Ci.is_cuda
        shapes = {"J": (nbf5, self.nbf, self.nbf), "K": (nbf5, self.nbf, self.nbf), "J_MO": (nbf5, nbf5),
                  "K_MO": (nbf5, nbf5), "H_core": (nbf5,)}
        res = {}
        for name in want:
            if out is not None and name in out:
                a = out[name]
                if tuple(a.shape) != shapes[name]:
                    raise ValueError(f"out[{name}] must have shape {shapes[name]}")
                if _is_torch(a):
                    if a.dtype != _torch().float64 or not a.is_contiguous():
                        raise ValueError(f"out[{name}] must be contiguous float64")
                elif a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError(f"out[{name}] must be C-contiguous float64")
                res[name] = a
            elif on_device:
                res[name] = _torch().empty(shapes[name], dtype=_torch().float64, device=Ci.device)
            else:
                res[name] = _host_empty(shapes[name])
        if nbf5 == 0:  # no orbitals: zero-size outputs, as the reference's slicing C[:, :0] gives
            return res
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(Ci.device).cuda_stream
        ptrs = [(_ptr(res[n]) if n in res else None) for n in ("J", "K", "J_MO", "K_MO", "H_core")]
        _lib.check(_lib.lib().pnjk_jk_build(self._h, _ptr(Ci), Ci.shape[1], nbf5,
                                            _ptr(Hi) if Hi is not None else None, self.nbf if Hi is not None else 0,
                                            prec, mask, *ptrs, stream))
        return res

    def mo_integrals(self, C, nbf5: int, H=None, precision: str = "fp64", outputs=("b_MO", "H_mat"), out=None,
                     stream=None):
        """Orbital-gradient MO integrals (JKH_MO_tmp with RI, integrals.py:665-716):
        b_MO (nbf, nbf5, naux) = C^T b C[:, :nbf5] per aux function and H_mat
        (nbf, nbf5) = C^T H C[:, :nbf5]. C: (nbf, nbf), every column used.
        precision: "fp64" or "split". Inputs and outputs follow `build`."""
        if self.nbf == 0:
            raise _lib.PnjkError(_lib.ESTATE, "no b uploaded")
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        nbf5 = int(nbf5)
        Ci = _as_input(C, "C")
        if Ci.ndim != 2 or tuple(Ci.shape) != (self.nbf, self.nbf):
            raise ValueError(f"C must have shape ({self.nbf}, {self.nbf}); got {tuple(Ci.shape)}")
        want = [n for n in outputs]
        for n in want:
            if n not in ("b_MO", "H_mat"):
                raise ValueError(f"unknown output {n!r}")
        Hi = None
        if "H_mat" in want:
            if H is None:
                raise ValueError("H_mat requested but H is None")
            Hi = _as_input(H, "H")
            if tuple(Hi.shape) != (self.nbf, self.nbf):
                raise ValueError(f"H must have shape ({self.nbf}, {self.nbf}); got {tuple(Hi.shape)}")
        on_device = _is_torch(Ci) and Ci.is_cuda
        shapes = {"b_MO": (self.nbf, nbf5, self.naux), "H_mat": (self.nbf, nbf5)}
        res = {}
        for name in want:
            if out is not None and name in out:
                a = out[name]
                if tuple(a.shape) != shapes[name]:
                    raise ValueError(f"out[{name}] must have shape {shapes[name]}")
                if _is_torch(a):
                    if a.dtype != _torch().float64 or not a.is_contiguous():
                        raise ValueError(f"out[{name}] must be contiguous float64")
                elif a.dtype != np.float64 or not a.flags.c_contiguous:
                    raise ValueError(f"out[{name}] must be C-contiguous float64")
                res[name] = a
            elif on_device:
                res[name] = _torch().empty(shapes[name], dtype=_torch().float64, device=Ci.device)
            else:
                res[name] = _host_empty(shapes[name])
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(Ci.device).cuda_stream
        _lib.check(_lib.lib().pnjk_mo_integrals(
            self._h, _ptr(Ci), Ci.shape[1], nbf5, _ptr(Hi) if Hi is not None else None,
            self.nbf if Hi is not None else 0, prec, _ptr(res["b_MO"]) if "b_MO" in res else None,
            _ptr(res["H_mat"]) if "H_mat" in res else None, stream))
        return res

    def lagrangian(self, C, nbf5: int, H, n, cj12, ck12, nbeta: int, nalpha: int, precision: str = "fp64",
                   want_hmat: bool = True, out=None, stream=None):
        """Closed-shell orbital-gradient Lagrangian (computeLagrange2 with RI, utils.py:262-355):
        returns {"elag": (nbf, nbf), "H_mat": (nbf, nbf5)}. n (nbf5,), cj12 and ck12
        (nbf5, nbf5) with diagonals taken as zero; inputs are not modified. b_MO stays on
        the device. precision "fp64" or "split" (of b_MO); contractions are FP64."""
        if self.nbf == 0:
            raise _lib.PnjkError(_lib.ESTATE, "no b uploaded")
        prec = _lib.PRECISIONS[precision] if isinstance(precision, str) else int(precision)
        nbf5 = int(nbf5)
        Ci, Hi = _as_input(C, "C"), _as_input(H, "H")
        if tuple(Ci.shape) != (self.nbf, self.nbf) or tuple(Hi.shape) != (self.nbf, self.nbf):
            raise ValueError(f"C and H must have shape ({self.nbf}, {self.nbf})")
        ni, cji, cki = _as_input(n, "n"), _as_input(cj12, "cj12"), _as_input(ck12, "ck12")
        if tuple(ni.shape) != (nbf5,) or tuple(cji.shape) != (nbf5, nbf5) or tuple(cki.shape) != (nbf5, nbf5):
            raise ValueError(f"n must have shape ({nbf5},), cj12 and ck12 ({nbf5}, {nbf5})")
        on_device = _is_torch(Ci) and Ci.is_cuda
        shapes = {"elag": (self.nbf, self.nbf), "H_mat": (self.nbf, nbf5)}
        res = {}
        for name in ("elag", "H_mat") if want_hmat else ("elag",):
            if out is not None and name in out:
                res[name] = out[name]
            elif on_device:
                res[name] = _torch().empty(shapes[name], dtype=_torch().float64, device=Ci.device)
            else:
                res[name] = _host_empty(shapes[name])
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(Ci.device).cuda_stream
        _lib.check(_lib.lib().pnjk_lagrangian(
            self._h, _ptr(Ci), Ci.shape[1], nbf5, _ptr(Hi), Hi.shape[1], _ptr(ni), _ptr(cji), _ptr(cki),
            int(nbeta), int(nalpha), prec, _ptr(res["elag"]), _ptr(res["H_mat"]) if "H_mat" in res else None,
            stream))
        return res

    def fock_rc(self, J, K, n, H, cj12, ck12, no1: int, nbeta: int, nalpha: int, out=None, stream=None):
        """Closed-shell generalised Fock matrices (computeF_RC_CPU/_GPU, utils.py:25-84):
        F (nbf5, nbf, nbf) from J, K (nbf5, nbf, nbf), n (nbf5,), H (nbf, nbf) and
        cj12, ck12 (nbf5, nbf5), whose [ini:, ini:] diagonals (ini = no1 if no1 > 1 else 0)
        are taken as zero; inputs are not modified. Needs no b. Device (torch) J gives a
        device F; host inputs give a host F."""
        return self._fock("pnjk_fock_rc", J, K, n, H, cj12, ck12, no1, nbeta, nalpha, out, stream)

    def fock_ro(self, J, K, n, H, cj12, ck12, no1: int, nbeta: int, nalpha: int, out=None, stream=None):
        """Restricted open-shell generalised Fock matrices (computeF_RO_CPU/_GPU,
        utils.py:96-246): as fock_rc, with the singly occupied orbitals [nbeta, nalpha)
        entering as the reference's SUMij and PRODWROij terms (utils.py:154-169) and the
        C^J/C^K sums restricted to the other orbitals."""
        return self._fock("pnjk_fock_ro", J, K, n, H, cj12, ck12, no1, nbeta, nalpha, out, stream)

    def _fock(self, fn, J, K, n, H, cj12, ck12, no1, nbeta, nalpha, out, stream):
        Ji, Ki, Hi = _as_input(J, "J"), _as_input(K, "K"), _as_input(H, "H")
        if Ji.ndim != 3 or Ji.shape[1] != Ji.shape[2] or tuple(Ki.shape) != tuple(Ji.shape):
            raise ValueError(f"J and K must have one shape (nbf5, nbf, nbf); got {tuple(Ji.shape)}, {tuple(Ki.shape)}")
        nbf5, nbf = int(Ji.shape[0]), int(Ji.shape[1])
        if tuple(Hi.shape) != (nbf, nbf):
            raise ValueError(f"H must have shape ({nbf}, {nbf}); got {tuple(Hi.shape)}")
        ni, cji, cki = _as_input(n, "n"), _as_input(cj12, "cj12"), _as_input(ck12, "ck12")
        if tuple(ni.shape) != (nbf5,) or tuple(cji.shape) != (nbf5, nbf5) or tuple(cki.shape) != (nbf5, nbf5):
            raise ValueError(f"n must have shape ({nbf5},), cj12 and ck12 ({nbf5}, {nbf5})")
        on_device = _is_torch(Ji) and Ji.is_cuda
        if out is not None:
            F = out
        elif on_device:
            F = _torch().empty((nbf5, nbf, nbf), dtype=_torch().float64, device=Ji.device)
        else:
            F = _host_empty((nbf5, nbf, nbf))
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(Ji.device).cuda_stream
        _lib.check(getattr(_lib.lib(), fn)(
            self._h, _ptr(Ji), _ptr(Ki), nbf, nbf5, _ptr(ni), _ptr(Hi), nbf, _ptr(cji), _ptr(cki), int(no1),
            int(nbeta), int(nalpha), _ptr(F), stream))
        return F

    # ---------------------------------------------------------------- occupation step
    @staticmethod
    def pnof(no1: int, ndoc: int, nsoc: int, ncwo: int, ipnof: int = 8, ista: int = 0,
             occ_method: str = "Softmax", HighSpin: bool = False, MSpin: int = 0):
        """struct pnjk_pnof from PyNOF's param fields (parameters.py:46-110)."""
        nbf5 = int(no1) + int(ndoc) * (1 + int(ncwo)) + int(nsoc)
        if occ_method not in _lib.OCC_METHODS:
            raise ValueError(f"occ_method must be one of {sorted(_lib.OCC_METHODS)}; got {occ_method!r}")
        return _lib.Pnof(int(ipnof), int(ista), _lib.OCC_METHODS[occ_method], int(bool(HighSpin)), int(MSpin),
                         int(no1), int(ndoc), int(nsoc), int(ncwo), nbf5)

    def occupation_step(self, gamma, pnof, J_MO=None, K_MO=None, H_core=None,
                        outputs=("E", "grad"), out=None, stream=None):
        """One occupation step (pnof.py:692-1025): any of E (calcocce), grad (calcoccg),
        n and dn_dgamma (ocupacion), cj12 and ck12 (PNOFi_selector) from the parameters
        gamma (nv,). E and grad need J_MO, K_MO (nbf5, nbf5) and H_core (nbf5,). Device
        (torch) inputs give device outputs; E is then a one-element tensor."""
        nbf5, nv = pnof.nbf5, pnof.ndoc * pnof.ncwo
        shapes = {"E": (1,), "grad": (nv,), "n": (nbf5,), "dn_dgamma": (nbf5, nv),
                  "cj12": (nbf5, nbf5), "ck12": (nbf5, nbf5)}
        for name in outputs:
            if name not in shapes:
                raise ValueError(f"unknown output {name!r}; choose from {list(shapes)}")
        gi = _as_input(gamma, "gamma") if nv > 0 else None
        if gi is not None and tuple(gi.shape) != (nv,):
            raise ValueError(f"gamma must have shape ({nv},); got {tuple(gi.shape)}")
        mats = [None, None, None]
        if "E" in outputs or "grad" in outputs:
            if J_MO is None or K_MO is None or H_core is None:
                raise ValueError("E and grad need J_MO, K_MO and H_core")
            mats = [_as_input(J_MO, "J_MO"), _as_input(K_MO, "K_MO"), _as_input(H_core, "H_core")]
            if tuple(mats[0].shape) != (nbf5, nbf5) or tuple(mats[1].shape) != (nbf5, nbf5) or \
                    tuple(mats[2].shape) != (nbf5,):
                raise ValueError(f"J_MO, K_MO must have shape ({nbf5}, {nbf5}) and H_core ({nbf5},)")
        ref = next((x for x in [gi] + mats if x is not None), None)
        on_device = ref is not None and _is_torch(ref) and ref.is_cuda
        res = {}
        for name in outputs:
            if out is not None and name in out:
                res[name] = out[name]
            elif on_device:
                res[name] = _torch().empty(shapes[name], dtype=_torch().float64, device=ref.device)
            else:
                res[name] = _host_empty(shapes[name])
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(ref.device).cuda_stream
        ptr = lambda x: None if x is None else _ptr(x)  # noqa: E731
        _lib.check(_lib.lib().pnjk_occupation_step(
            self._h, ctypes.addressof(pnof), ptr(gi), ptr(mats[0]), ptr(mats[1]), ptr(mats[2]),
            *(ptr(res.get(k)) for k in ("E", "grad", "n", "dn_dgamma", "cj12", "ck12")), stream))
        return res

    def pnof_coefficients(self, n, pnof, stream=None):
        """(cj12, ck12) for occupations n (PNOFi_selector, pnof.py:657-667)."""
        ni = _as_input(n, "n")
        if tuple(ni.shape) != (pnof.nbf5,):
            raise ValueError(f"n must have shape ({pnof.nbf5},); got {tuple(ni.shape)}")
        shape = (pnof.nbf5, pnof.nbf5)
        if _is_torch(ni) and ni.is_cuda:
            cj = _torch().empty(shape, dtype=_torch().float64, device=ni.device)
            ck = _torch().empty_like(cj)
            stream = stream if stream is not None else _torch().cuda.current_stream(ni.device).cuda_stream
        else:
            cj, ck = np.empty(shape), np.empty(shape)
        _lib.check(_lib.lib().pnjk_pnof_coefficients(self._h, ctypes.addressof(pnof), _ptr(ni), _ptr(cj), _ptr(ck),
                                                      stream))
        return cj, ck

    def energy(self, n, cj12, ck12, J_MO, K_MO, H_core, nbeta: int, nalpha: int, MSpin: int = 0,
               stream=None) -> float:
        """Energy for given coefficients (calce, pnof.py:905-990; MSpin != 0 the open-shell form)."""
        arrs = [_as_input(x, nm) for x, nm in ((n, "n"), (cj12, "cj12"), (ck12, "ck12"), (J_MO, "J_MO"),
                                               (K_MO, "K_MO"), (H_core, "H_core"))]
        nbf5 = int(arrs[0].shape[0])
        for a, shp in zip(arrs, [(nbf5,)] + [(nbf5, nbf5)] * 4 + [(nbf5,)]):
            if tuple(a.shape) != shp:
                raise ValueError(f"shapes must be n ({nbf5},), matrices ({nbf5}, {nbf5}), H_core ({nbf5},)")
        E = np.empty(1)
        if stream is None and _is_torch(arrs[0]) and arrs[0].is_cuda:
            stream = _torch().cuda.current_stream(arrs[0].device).cuda_stream
        _lib.check(_lib.lib().pnjk_energy(self._h, nbf5, int(nbeta), int(nalpha), int(MSpin),
                                          *(_ptr(a) for a in arrs), E.ctypes.data, stream))
        return float(E[0])

    def lagrange_convergence(self, elag, stream=None):
        """(sumdiff, maxdiff) = (sum |elag - elag^T|, max |elag - elag^T|), computeLagrangeConvergency
        (utils.py:405-411); elag host or device (nbf, nbf)."""
        ei = _as_input(elag, "elag")
        if ei.ndim != 2 or ei.shape[0] != ei.shape[1]:
            raise ValueError(f"elag must be square; got {tuple(ei.shape)}")
        if stream is None and _is_torch(ei) and ei.is_cuda:
            stream = _torch().cuda.current_stream(ei.device).cuda_stream
        sd, md = ctypes.c_double(), ctypes.c_double()
        _lib.check(_lib.lib().pnjk_lagrange_convergence(self._h, _ptr(ei), int(ei.shape[0]), ctypes.byref(sd),
                                                         ctypes.byref(md), stream))
        return sd.value, md.value

    def lagrange_fock(self, F, C, noptorb: int, out=None, stream=None):
        """Lagrange multipliers from the Fock matrices (computeLagrange, utils.py:249-259):
        elag (nbf, nbf) with elag[i, j] = sum_mn C[m, i] F[j, m, n] C[n, j] for
        i < noptorb, j < nbf5, zero elsewhere. F (nbf5, nbf, nbf), C (nbf, nbf). Needs no b."""
        Fi, Ci = _as_input(F, "F"), _as_input(C, "C")
        if Fi.ndim != 3 or Fi.shape[1] != Fi.shape[2]:
            raise ValueError(f"F must have shape (nbf5, nbf, nbf); got {tuple(Fi.shape)}")
        nbf5, nbf = int(Fi.shape[0]), int(Fi.shape[1])
        if tuple(Ci.shape) != (nbf, nbf):
            raise ValueError(f"C must have shape ({nbf}, {nbf}); got {tuple(Ci.shape)}")
        on_device = _is_torch(Fi) and Fi.is_cuda
        if out is not None:
            elag = out
        elif on_device:
            elag = _torch().empty((nbf, nbf), dtype=_torch().float64, device=Fi.device)
        else:
            elag = np.empty((nbf, nbf), dtype=np.float64)
        if stream is None and on_device:
            stream = _torch().cuda.current_stream(Fi.device).cuda_stream
        _lib.check(_lib.lib().pnjk_lagrange_fock(self._h, _ptr(Fi), nbf, nbf5, _ptr(Ci), nbf, int(noptorb),
                                                  _ptr(elag), stream))
        return elag
