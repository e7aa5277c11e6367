"""Tile shapes change no result bit (GPU only): every stage, every supported
shape, FP64 and split forms, against the built-in choice."""
import numpy as np
import pytest

from pynof_amd import JKEngine, _lib, autotune, synth

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", ["fp64", "split"])
def test_tiles_bitwise(gpu_required, precision):
    nbf, naux, nbf5 = 150, 300, 110
    print(f"case: tiles bitwise {precision} nbf={nbf} naux={naux} nbf5={nbf5}", flush=True)
    b = synth.make_b(nbf, naux, 77, 0.05)
    C = synth.make_C(nbf, 78)
    H = synth.make_H(nbf, 79)
    eng = JKEngine(0)
    eng.upload_b(b)
    ref = eng.build(C, nbf5, H=H, precision=precision)
    L = _lib.lib()
    a = autotune.ARITH[precision]
    for st, idx in autotune.STAGES.items():
        for fm, fn in autotune.CANDIDATES[st]:
            if precision == "fp64" and (fm, fn) in ((10, 4), (5, 5)):
                assert L.pnjk_set_tile(eng._h, a, idx, fm, fn) == _lib.EINVAL  # BF16x3-only shapes
                continue
            _lib.check(L.pnjk_set_tile(eng._h, a, idx, fm, fn))
            r = eng.build(C, nbf5, H=H, precision=precision)
            for k in ref:
                assert np.array_equal(r[k], ref[k]), (st, fm, fn, k)
        _lib.check(L.pnjk_set_tile(eng._h, a, idx, 0, 0))
    eng.close()


def test_set_tile_rejects(gpu_required):
    eng = JKEngine(0)
    L = _lib.lib()
    assert L.pnjk_set_tile(eng._h, 0, 2, 5, 1) == _lib.EINVAL   # not an offered shape
    assert L.pnjk_set_tile(eng._h, 0, 2, 4, 4) == _lib.OK       # 128-column tiles, FP64 and BF16x3
    assert L.pnjk_set_tile(eng._h, 1, 2, 4, 4) == _lib.OK
    assert L.pnjk_set_tile(eng._h, 1, 2, 0, 0) == _lib.OK
    assert L.pnjk_set_tile(eng._h, 1, 1, 10, 4) == _lib.OK      # eight-wave tile: BF16x3 HT and J^n
    assert L.pnjk_set_tile(eng._h, 1, 6, 10, 4) == _lib.EINVAL  # not the exchange SYRK
    assert L.pnjk_set_tile(eng._h, 1, 2, 10, 4) == _lib.EINVAL  # nor the quarter transform
    assert L.pnjk_set_tile(eng._h, 1, 1, 5, 5) == _lib.OK       # 160 x 160 tile: BF16x3 HT and J^n
    assert L.pnjk_set_tile(eng._h, 0, 1, 5, 5) == _lib.EINVAL
    assert L.pnjk_set_tile(eng._h, 1, 6, 5, 5) == _lib.OK       # and the BF16x3 exchange SYRK
    assert L.pnjk_set_tile(eng._h, 0, 6, 5, 5) == _lib.EINVAL
    assert L.pnjk_set_tile(eng._h, 1, 6, 5, 4) == _lib.EINVAL   # exchange SYRK needs square tiles
    assert L.pnjk_set_tile(eng._h, 0, 0, 2, 2) == _lib.EINVAL   # copy_in takes no tile
    assert L.pnjk_set_tile(eng._h, 1, 1, 3, 3) == _lib.EINVAL   # unknown shape
    assert L.pnjk_set_tile(eng._h, 2, 1, 2, 2) == _lib.EINVAL   # unknown arithmetic
    eng.close()
