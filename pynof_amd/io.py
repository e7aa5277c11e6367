"""On-disk formats for the inputs of a build: b, C, n (and H).

The reference keeps b, C and n in memory only (b from psi4 at integrals.py:35-53;
C and n in the optimiser's state). Here:

* b is an ``.npy`` file in the reference layout (nbf, nbf, naux), float64,
  C-order. ``load_b`` memory-maps it, and ``JKEngine.upload_b`` streams a
  memory-mapped array to the device in mu-row chunks, so a 28.8 GB b never
  needs to fit in host memory at once. ``save_b_from_engine`` writes a resident
  b back through a memory-mapped file in the same way.
* C, n, H and the sizes are one ``.npz`` (``save_state`` / ``load_state``),
  read without pickle.
"""
from __future__ import annotations

import os

import numpy as np

_STATE_KEYS = ("C", "n", "H", "gamma")


def save_b(path: str | os.PathLike, b) -> None:
    """Writes b (nbf, nbf, naux) as float64 .npy."""
    b = np.asarray(b)
    if b.ndim != 3 or b.shape[0] != b.shape[1]:
        raise ValueError(f"b must have shape (nbf, nbf, naux); got {b.shape}")
    np.save(path, np.ascontiguousarray(b, dtype=np.float64), allow_pickle=False)


def load_b(path: str | os.PathLike, mmap: bool = True) -> np.ndarray:
    """b from a .npy file, memory-mapped read-only by default."""
    b = np.load(path, mmap_mode="r" if mmap else None, allow_pickle=False)
    if b.dtype != np.float64 or b.ndim != 3 or b.shape[0] != b.shape[1] or not b.flags.c_contiguous:
        raise ValueError(f"{path}: expected a C-order float64 (nbf, nbf, naux) array; got {b.dtype} {b.shape}")
    return b


def save_b_from_engine(path: str | os.PathLike, engine) -> None:
    """Writes the engine's resident b to a .npy file through a memory map (no full host copy)."""
    if engine.nbf == 0:
        raise ValueError("the engine holds no b")
    shape = (engine.nbf, engine.nbf, engine.naux)
    out = np.lib.format.open_memmap(path, mode="w+", dtype=np.float64, shape=shape)
    engine.download_b(out)
    out.flush()
    del out


def save_state(path: str | os.PathLike, *, nbf5: int, C=None, n=None, H=None, gamma=None, **sizes) -> None:
    """C (nbf, nbf), n (nbf5,), H (nbf, nbf), gamma and integer sizes (nbf5, ndoc, ...) in one .npz."""
    arrays = {"nbf5": np.int64(nbf5)}
    for key, val in (("C", C), ("n", n), ("H", H), ("gamma", gamma)):
        if val is not None:
            arrays[key] = np.ascontiguousarray(np.asarray(val, dtype=np.float64))
    for key, val in sizes.items():
        if not isinstance(val, (int, np.integer)):
            raise TypeError(f"size {key} must be an integer")
        arrays[key] = np.int64(val)
    np.savez(path, **arrays)


def load_state(path: str | os.PathLike) -> dict:
    """The arrays and sizes written by save_state (no pickle)."""
    with np.load(path, allow_pickle=False) as z:
        out = {}
        for key in z.files:
            v = z[key]
            out[key] = int(v) if v.ndim == 0 else v
    for key in _STATE_KEYS:
        if key in out and out[key].dtype != np.float64:
            raise ValueError(f"{path}: {key} is {out[key].dtype}, expected float64")
    return out
