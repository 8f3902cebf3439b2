"""CPU: the Python bindings check buffer sizes and shapes before they hand raw
pointers to the library (no device call is reached on these paths)."""
import ctypes as C

import numpy as np
import pytest

import bec_amd as B
from bec_amd import fno as F


def _fake_device_array(nbytes):
    d = F.DeviceArray.__new__(F.DeviceArray)
    d.shape, d.dtype, d.nbytes = (nbytes // 4,), np.dtype(np.float32), nbytes
    d.ptr = C.c_void_p(0x10000)
    return d


def _calc():
    return B.derive(B.default_model(N=30000, nvx=16, nvy=16, nvz=128, ne=1))


def test_gs2_features_rejects_short_buffer():
    c = _calc()
    short = _fake_device_array(c.cells * 64 - 4)
    with pytest.raises(ValueError, match="need"):
        F.gs2_features(c, [0.0, 0.0, 0.0, 0.0], short)
    # a view that starts inside the allocation has only the rest of it
    base = _fake_device_array(2 * c.cells * 64)
    view = F.DeviceView(base, c.cells * 64 + 4)
    with pytest.raises(ValueError, match="need"):
        F.gs2_features(c, [0.0, 0.0, 0.0, 0.0], view)
    with pytest.raises(ValueError):
        F.gs2_features(c, [0.0, 0.0, 0.0], _fake_device_array(c.cells * 64))
    short.ptr = base.ptr = C.c_void_p()   # nothing to free


def test_set_state_rejects_wrong_shape():
    s = B.Solver.__new__(B.Solver)
    s.groups = [_calc()]
    s._ctx = C.c_void_p()
    good = np.zeros(s.shape(0), dtype=complex)
    with pytest.raises(ValueError, match="shape"):
        s.set_state(good[:, :8], good)
    with pytest.raises(ValueError, match="shape"):
        s.set_state(good, good.reshape(-1))
