"""Tile-shape autotuner for the JK build stages.

For one workload shape and precision it times every supported tile shape of
each tunable stage (half_transform, quarter_transform, coulomb_jn, exchange_kn)
with the engine's HIP-event stage timer, and keeps the fastest. Every output
element's k-summation order is fixed by the MFMA sequence, so a tile shape
changes no result bit (tests/test_gpu_tiles.py checks this); the tuned table is
therefore free to differ between batch mixes without changing results.

Tables are JSON: {"<precision>:<nbf>:<naux>:<nbf5>": {"<stage>": [fm, fn], ...}}.
The "fp64" entry tunes the FP64 kernels, the "split" entry the BF16x3 kernels
(the split_refine form uses the FP64 half-transform choice for its FP64
half-transform, triangular or full, a fixed tile for its FP64 MO transform or the
FP64 quarter-transform choice, and the split choices for its BF16x3
half-transform, J^n and K^n).
``lookup`` returns the entry of the nearest tuned shape (log-distance in nbf,
naux, nbf5), so instances of untuned sizes get the choice of their nearest
tuned neighbour.
"""
from __future__ import annotations

import ctypes
import json
import math
import os

from . import _lib

STAGES = {"half_transform": 1, "quarter_transform": 2, "coulomb_jn": 5, "exchange_kn": 6}
CANDIDATES = {
    "half_transform": [(2, 2), (4, 4), (5, 4), (4, 2), (2, 4), (5, 2), (10, 4), (5, 5)],  # last two: BF16x3 only
    "quarter_transform": [(2, 2), (4, 2), (5, 2), (2, 4), (4, 4), (5, 4)],
    "coulomb_jn": [(2, 2), (4, 4), (5, 4), (4, 2), (2, 4), (5, 2), (10, 4), (5, 5)],
    "exchange_kn": [(2, 2), (4, 4), (5, 5)],  # (5, 5): BF16x3 only
}
DEFAULT_TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_tiles.json")


def key(precision: str, nbf: int, naux: int, nbf5: int) -> str:
    return f"{precision}:{nbf}:{naux}:{nbf5}"


def stage_times(eng, build, reps: int = 2):
    """Runs `build()` reps times with stage profiling on; returns the min time per stage (ms)."""
    L = _lib.lib()
    n = L.pnjk_stage_count()
    names = [L.pnjk_stage_name(i).decode() for i in range(n)]
    buf = (ctypes.c_double * n)()
    best = {nm: float("inf") for nm in names}
    _lib.check(L.pnjk_set_profiling(eng._h, 1))
    try:
        for _ in range(reps):
            build()
            _lib.check(L.pnjk_stage_times(eng._h, buf))
            for i, nm in enumerate(names):
                best[nm] = min(best[nm], buf[i])
    finally:
        _lib.check(L.pnjk_set_profiling(eng._h, 0))
    return best


ARITH = {"fp64": 0, "split": 1}


def tune(eng, build, arith: str, reps: int = 2, stages=None, log=None):
    """Tunes the tile of each stage of the `arith` kernels ("fp64" or "split") in
    turn, others at their current choice. `build()` runs one build on `eng` in a
    form that uses those kernels. Returns {stage: [fm, fn]}."""
    L = _lib.lib()
    a = ARITH[arith]
    chosen = {}
    for st in stages or STAGES:
        idx = STAGES[st]
        timings = {}
        for fm, fn in CANDIDATES[st]:
            if L.pnjk_set_tile(eng._h, a, idx, fm, fn) != _lib.OK:  # shape not offered by this arithmetic
                continue
            build()  # warm the shape's code object and workspace
            timings[(fm, fn)] = stage_times(eng, build, reps)[st]
        best = min(timings, key=timings.get)
        _lib.check(L.pnjk_set_tile(eng._h, a, idx, *best))
        chosen[st] = list(best)
        if log:
            log(f"{st}: " + ", ".join(f"{k}={v:.3f}ms" for k, v in sorted(timings.items())) + f" -> {best}")
    return chosen


_loaded: dict = {}  # path -> (modification time, table): groups are created often (serving)


def load(path: str = DEFAULT_TABLE) -> dict:
    """The table at `path` ({} if absent); parsed again only when the file changes.
    Returns a fresh top-level dict; treat the entries as read-only."""
    try:
        mtime = os.stat(path).st_mtime_ns
    except FileNotFoundError:
        return {}
    hit = _loaded.get(path)
    if hit is None or hit[0] != mtime:
        with open(path) as f:
            hit = _loaded[path] = (mtime, json.load(f))
    return dict(hit[1])


def save(table: dict, path: str = DEFAULT_TABLE) -> None:
    with open(path, "w") as f:
        json.dump(table, f, indent=1, sort_keys=True)
    _loaded.pop(path, None)


def lookup(table: dict, precision: str, nbf: int, naux: int, nbf5: int):
    """Entry of the nearest tuned shape of the same precision, or None."""
    best, dist = None, float("inf")
    for k, v in table.items():
        p, a, b_, c = k.split(":")
        if p != precision:
            continue
        d = (math.log(int(a) / nbf) ** 2 + math.log(int(b_) / max(naux, 1)) ** 2
             + math.log(int(c) / nbf5) ** 2)
        if d < dist:
            best, dist = v, d
    return best


def apply(eng, arith: str, entry) -> None:
    """Sets the engine's stage tiles of the `arith` kernels from a table entry
    (None restores the built-in choices)."""
    L = _lib.lib()
    for st, idx in STAGES.items():
        fm, fn = (entry or {}).get(st, (0, 0))
        _lib.check(L.pnjk_set_tile(eng._h, ARITH[arith], idx, int(fm), int(fn)))


def tune_sizes(sizes, out_path: str, reps: int = 2, log=print) -> dict:
    """Tunes both arithmetics at each (nbf, naux, nbf5) with device-drawn synthetic
    b and seeded C, H; writes the table to out_path and returns it."""
    import torch

    from .engine import JKEngine
    from .synth import device_b_packed, make_C, make_H

    table = {}
    for nbf, naux, nbf5 in sizes:
        eng = JKEngine(0)
        bt = device_b_packed(nbf, naux, 4242, 0.05, JKEngine.packed_ld(nbf, naux))
        eng.upload_b_packed(bt, nbf, naux)
        del bt
        C = torch.from_numpy(make_C(nbf, 4243)).cuda()
        H = torch.from_numpy(make_H(nbf, 4244)).cuda()
        for arith, form in (("split", "split"), ("fp64", "fp64")):
            out = None

            def build():
                nonlocal out
                out = eng.build(C, nbf5, H=H, precision=form, out=out)

            build()
            table[key(arith, nbf, naux, nbf5)] = tune(eng, build, arith, reps=reps,
                                                      log=lambda m: log(f"  {arith} {nbf}/{naux}/{nbf5} {m}"))
            out = None
        eng.close()
        torch.cuda.empty_cache()
        save(table, out_path)
    return table


if __name__ == "__main__":
    import sys

    from .synth import ceil8

    path = sys.argv[1] if len(sys.argv) > 1 else DEFAULT_TABLE
    grid = [(n, ceil8(3.6 * n), n) for n in (48, 96, 160, 240, 320, 400, 560, 800)]
    grid += [(280, 1100, 280), (1000, 3600, 300)]
    tune_sizes(grid, path, log=lambda m: print(m, flush=True))
