"""On-disk formats (pynof_amd/io.py): round trips on the CPU."""
import numpy as np
import pytest

from pynof_amd import io, synth


def test_b_round_trip_memmap(tmp_path):
    b = synth.make_b(7, 9, 1234, 0.05)
    path = tmp_path / "b.npy"
    io.save_b(path, b)
    m = io.load_b(path)
    assert isinstance(m, np.memmap) and np.array_equal(m, b)
    assert np.array_equal(io.load_b(path, mmap=False), b)


def test_b_rejects_wrong_shape_and_dtype(tmp_path):
    with pytest.raises(ValueError):
        io.save_b(tmp_path / "x.npy", np.zeros((3, 4, 5)))
    np.save(tmp_path / "f32.npy", np.zeros((3, 3, 2), dtype=np.float32))
    with pytest.raises(ValueError):
        io.load_b(tmp_path / "f32.npy")


def test_state_round_trip(tmp_path):
    C = synth.make_C(6, 1)
    n = synth.make_n(6, 2, 2)
    H = synth.make_H(6, 3)
    path = tmp_path / "state.npz"
    io.save_state(path, nbf5=6, C=C, n=n, H=H, ndoc=2, ncwo=2)
    s = io.load_state(path)
    assert s["nbf5"] == 6 and s["ndoc"] == 2 and s["ncwo"] == 2
    assert np.array_equal(s["C"], C) and np.array_equal(s["n"], n) and np.array_equal(s["H"], H)
    assert "gamma" not in s
    with pytest.raises(TypeError):
        io.save_state(path, nbf5=6, ndoc=2.5)
