"""The allocation-free forward of the corrector's network
(mlc_mlp_forward_noalloc, scratch in a static device array) against the
allocating mlc_mlp_forward: the same bits on two calls in a row, the second
with fewer rows than the first (the scratch keeps the first call's
contents), and a clear refusal of more rows than the scratch holds.  Device
buffers for X and Y come from the HIP runtime through ctypes."""
from __future__ import annotations

import ctypes as C

import numpy as np
import pytest

import dynamical_model_amd as dm
from dynamical_model_amd import _lib
from dynamical_model_amd.lac import MLP, nin_for, nout_for

pytestmark = pytest.mark.gpu


def _hip():
    try:
        h = C.CDLL("libamdhip64.so.7")
    except OSError:
        h = C.CDLL("libamdhip64.so")
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    return h


def _dev(h, nbytes):
    p = C.c_void_p()
    assert h.hipMalloc(C.byref(p), nbytes) == 0
    return p.value


def test_noalloc_forward_matches_allocating_forward():
    kz = 40  # the corrector's network at 512 x 512 x 40: 2160 -> 512 -> 512 -> 160
    nin, nhid, nout = nin_for(kz), 512, nout_for(kz)
    net = MLP(nin, nhid, nout, seed=7)
    rng = np.random.default_rng(5)
    for layer in range(3):
        w, b = net.get_weights(layer)
        net.set_weights(layer, w, rng.standard_normal(b.shape).astype(np.float32) * 0.1)
    h = _hip()
    rows_max = 3000
    dx = _dev(h, rows_max * nin * 4)
    dy = _dev(h, rows_max * nout * 4)
    try:
        for rows in (rows_max, 700):
            print(f"noalloc forward {nin}->{nhid}->{nhid}->{nout}, {rows} rows", flush=True)
            x = rng.standard_normal((rows, nin)).astype(np.float32)
            ref = net.forward(x)
            assert h.hipMemcpy(C.c_void_p(dx), x.ctypes.data_as(C.c_void_p), x.nbytes, 1) == 0
            net.forward_noalloc(dx, rows, dy)
            y = np.zeros((rows, nout), np.float32)
            assert h.hipMemcpy(y.ctypes.data_as(C.c_void_p), C.c_void_p(dy), y.nbytes, 2) == 0
            assert np.abs(ref).max() > 0
            assert np.array_equal(y, ref)
        # more rows than the scratch holds: refused before any device work
        with pytest.raises(dm.MolochError) as e:
            net.forward_noalloc(dx, 20000, dy)
        assert e.value.code == _lib.MLC_EINVAL
        assert "scratch" in str(e.value)
    finally:
        h.hipFree(C.c_void_p(dx))
        h.hipFree(C.c_void_p(dy))
        net.close()
