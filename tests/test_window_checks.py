"""Host-side checks of the differentiable window's backward pass (no device):
a backward pass must belong to the lattice's last forward window, and its
cotangents must have that window's shapes.  The library call is never
reached when a check fails."""
import types

import numpy as np
import pytest
import torch

import lbdem_psm as L
from lbdem_psm.autograd import PSMWindow


class _NoCalls:
    def __getattr__(self, name):
        raise AssertionError(f"library call {name} reached")


def _lattice(window, nc=10):
    lat = object.__new__(L.Lattice)
    lat.lib = _NoCalls()
    lat.h = None
    lat.nc = nc
    lat.np_ = window[2] if window else 0
    lat._window = window
    return lat


def test_backward_needs_a_forward_window():
    with pytest.raises(L.PSMError) as e:
        _lattice(None).diff_backward(np.zeros((1, 1, 6)))
    assert e.value.code == L.ERR_STATE


def test_backward_checks_the_cotangent_shapes():
    lat = _lattice((1, 5, 3))
    with pytest.raises(L.PSMError) as e:
        lat.diff_backward(np.zeros((4, 3, 6)))     # a 4-step window's cotangent
    assert e.value.code == L.ERR_INVALID_ARG
    with pytest.raises(L.PSMError) as e:
        lat.diff_backward(np.zeros((5, 2, 6)))
    assert e.value.code == L.ERR_INVALID_ARG
    with pytest.raises(L.PSMError) as e:
        lat.diff_backward(None, np.zeros((19, 9)))  # 19 x Nc with Nc = 10 expected
    assert e.value.code == L.ERR_INVALID_ARG


def test_backward_refuses_a_superseded_window():
    lat = _lattice((2, 5, 3))
    with pytest.raises(L.PSMError) as e:
        lat.diff_backward(np.zeros((5, 3, 6)), token=1)
    assert e.value.code == L.ERR_STATE
    ctx = types.SimpleNamespace(lattice=lat, token=1, dtype=torch.float64)
    with pytest.raises(L.PSMError):
        PSMWindow.backward(ctx, torch.zeros(5, 3, 6, dtype=torch.float64))
