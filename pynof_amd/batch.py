"""Grouped JK builds: many device-resident instances in one pass.

``JKGroup`` wraps ``pnjk_group``: each GEMM stage of every instance runs in one
grouped launch (one workgroup per tile) per problem type and tile shape, optionally replayed
from a captured hipGraph. Tile shapes come from the autotuner's table
(``autotune.lookup``) per instance size; results are bitwise those of building
each instance alone (tile shapes do not change results).
"""
from __future__ import annotations

import ctypes

from . import _lib, autotune
from .engine import JKEngine, OUTPUT_BITS, _torch

_ORDER = ("J", "K", "J_MO", "K_MO", "H_core")


def _ptr_array(ptrs):
    arr = (ctypes.c_void_p * len(ptrs))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr


class JKGroup:
    """A fixed set of engines built together."""

    def __init__(self, engines: list[JKEngine], tuned: dict | None = None, precision_hint: str = "split",
                 nbf5s=None):
        """nbf5s: the orbital counts the group will build with, when known: their tiles are
        applied to the engines now (otherwise at the first build), so single builds on the
        same engines run the same tiles."""
        if not engines:
            raise ValueError("a group needs at least one engine")
        self.engines = list(engines)
        self._table = autotune.load() if tuned is None else tuned
        self._tiled_for = None  # (nbf, naux, nbf5) per engine of the tiles last applied
        if nbf5s is not None:
            self._apply_tiles(nbf5s)
        L = _lib.lib()
        handles = _ptr_array([e._h.value for e in self.engines])
        h = ctypes.c_void_p()
        _lib.check(L.pnjk_group_create(ctypes.cast(handles, ctypes.POINTER(ctypes.c_void_p)),
                                       len(self.engines), ctypes.byref(h)))
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            _lib.lib().pnjk_group_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _apply_tiles(self, nbf5s):
        """Tile shapes of each engine from the tuned table entry nearest its (nbf, naux,
        nbf5), looked up with the nbf5 of this build (tiles never change results)."""
        shapes = tuple((e.nbf, e.naux, int(n5)) for e, n5 in zip(self.engines, nbf5s))
        if shapes == self._tiled_for:
            return
        for e, (nbf, naux, n5) in zip(self.engines, shapes):
            for arith in ("fp64", "split"):
                entry = autotune.lookup(self._table, arith, nbf, naux, n5) if self._table else None
                autotune.apply(e, arith, entry)
        self._tiled_for = shapes

    def alloc_outputs(self, nbf5s, outputs=_ORDER, device=None):
        torch = _torch()
        outs = []
        for e, n5 in zip(self.engines, nbf5s):
            dev = device or torch.device("cuda", e.device)
            shapes = {"J": (n5, e.nbf, e.nbf), "K": (n5, e.nbf, e.nbf), "J_MO": (n5, n5), "K_MO": (n5, n5),
                      "H_core": (n5,)}
            outs.append({k: torch.empty(shapes[k], dtype=torch.float64, device=dev) for k in outputs})
        return outs

    def build(self, Cs, nbf5s, Hs=None, precision: str = "fp64", outputs=_ORDER, outs=None,
              use_graph: bool = False, stream=None):
        """Cs[i]: device (nbf_i, >= nbf5_i) FP64 tensors; outs: list of dicts of device
        tensors (allocated when None). Asynchronous on `stream` (torch's current stream
        by default). Returns the output dicts."""
        torch = _torch()
        n = len(self.engines)
        if len(Cs) != n or len(nbf5s) != n:
            raise ValueError("one C and one nbf5 per engine")
        mask = 0
        for k in outputs:
            mask |= OUTPUT_BITS[k]
        if "H_core" in outputs and Hs is None:
            raise ValueError("H_core requested but Hs is None")
        Cs = [c.contiguous() for c in Cs]
        Hs = [h.contiguous() for h in Hs] if Hs is not None else None
        for e, c, n5 in zip(self.engines, Cs, nbf5s):
            if c.dtype != torch.float64 or not c.is_cuda or c.shape[0] != e.nbf or c.shape[1] < n5:
                raise ValueError("each C must be a float64 device tensor of shape (nbf, >= nbf5)")
        self._apply_tiles(nbf5s)
        if outs is None:
            outs = self.alloc_outputs(nbf5s, outputs)
        prec = _lib.PRECISIONS[precision]
        i64 = ctypes.c_int64 * n
        arrs = {}
        for k in _ORDER:
            arrs[k] = _ptr_array([o[k].data_ptr() for o in outs]) if k in outputs else None
        if stream is None:
            stream = torch.cuda.current_stream(Cs[0].device).cuda_stream
        self._keep = (Cs, Hs, outs)  # inputs/outputs must outlive the asynchronous build
        _lib.check(_lib.lib().pnjk_group_build(
            self._h, _ptr_array([c.data_ptr() for c in Cs]), i64(*[c.shape[1] for c in Cs]), i64(*nbf5s),
            _ptr_array([h.data_ptr() for h in Hs]) if Hs is not None else None,
            i64(*[h.shape[1] for h in Hs]) if Hs is not None else None,
            prec, mask, arrs["J"], arrs["K"], arrs["J_MO"], arrs["K_MO"], arrs["H_core"], int(use_graph), stream))
        return outs
