"""Host-side argument checks of the engine's Python mirror (no device)."""
import numpy as np
import pytest

from bayadera_amd import engine


def test_out_accepts_a_matching_buffer():
    buf = np.zeros(12, np.float32)
    res, arr = engine._out(buf, (12,))
    assert res is buf and arr.ptr == buf.ctypes.data


def test_out_allocates_when_absent():
    res, arr = engine._out(None, (3, 4))
    assert res.shape == (3, 4) and res.dtype == np.float32 and arr.ptr == res.ctypes.data


def test_out_rejects_wrong_dtype():
    with pytest.raises(ValueError, match="dtype"):
        engine._out(np.zeros(12, np.float64), (12,))


def test_out_rejects_non_contiguous():
    with pytest.raises(ValueError, match="contiguous"):
        engine._out(np.zeros(24, np.float32)[::2], (12,))


def test_out_rejects_short_buffer():
    with pytest.raises(ValueError, match="holds 11"):
        engine._out(np.zeros(11, np.float32), (12,))


def test_out_rejects_other_types():
    with pytest.raises(ValueError):
        engine._out([0.0] * 12, (12,))
