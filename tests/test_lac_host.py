"""Host-side argument checks of the corrector's Python wrappers (no GPU):
arrays whose shapes do not match the network are refused before any call
into the library, which would otherwise read or write past them."""
from __future__ import annotations

import numpy as np
import pytest

from dynamical_model_amd.lac import MLP, nin_for, nout_for


def _net():
    # a borrowed handle is never dereferenced by the checks below
    return MLP(nin_for(8), 64, nout_for(8), _borrowed=1)


def test_network_dimensions():
    n = _net()
    assert (n.nin, n.nhid, n.nout) == (9 * 48, 64, 32)
    assert n.dims(0) == (432, 64) and n.dims(1) == (64, 64) and n.dims(2) == (64, 32)


@pytest.mark.parametrize("layer,w,b", [(0, (431, 64), (64,)), (1, (64, 64), (63,)), (2, (64, 33), (33,)),
                                       (2, (32, 64), (64,))])
def test_set_weights_rejects_wrong_shapes(layer, w, b):
    with pytest.raises(ValueError):
        _net().set_weights(layer, np.zeros(w, np.float32), np.zeros(b, np.float32))


def test_layer_out_of_range():
    n = _net()
    for bad in (-1, 3):
        with pytest.raises(ValueError):
            n.get_weights(bad)
        with pytest.raises(ValueError):
            n.gradients(bad)


def test_forward_and_grad_reject_wrong_shapes():
    n = _net()
    with pytest.raises(ValueError):
        n.forward(np.zeros((4, n.nin + 1), np.float32))
    with pytest.raises(ValueError):
        n.forward(np.zeros(n.nin, np.float32))
    with pytest.raises(ValueError):
        n.grad(np.zeros((4, n.nin), np.float32), np.zeros((4, n.nout - 1), np.float32))
    with pytest.raises(ValueError):
        n.grad(np.zeros((4, n.nin), np.float32), np.zeros((5, n.nout), np.float32))
