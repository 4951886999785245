"""Learned acoustic corrector: the column-stencil MLP engine (C ABI mlc_mlp_*).

The network maps a 3 x 3 stencil of columns -- u, v, w, Exner, theta, qv at
kz levels each -- to corrections of u, v, w and Exner at kz levels, trained
on pairs generated by the plain GPU core.  The reference has no learned
component; this is new.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load


def nin_for(kz: int) -> int:
    return 9 * 6 * kz


def nout_for(kz: int) -> int:
    return 4 * kz


class MLP:
    """nin -> nhid -> nhid -> nout with ReLU, BF16 MFMA, FP32 master weights."""

    def __init__(self, nin: int, nhid: int, nout: int, seed: int = 1, device: int = 0):
        self.nin, self.nhid, self.nout = nin, nhid, nout
        h = C.c_void_p()
        check("mlc_mlp_create", load().mlc_mlp_create(nin, nhid, nout, seed, device, C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def dims(self, layer: int) -> tuple[int, int]:
        return ([self.nin, self.nhid, self.nhid][layer], [self.nhid, self.nhid, self.nout][layer])

    def get_weights(self, layer: int):
        i, o = self.dims(layer)
        w = np.zeros((i, o), np.float32)
        b = np.zeros(o, np.float32)
        check("mlc_mlp_weights", load().mlc_mlp_weights(self._h, layer, w.ctypes.data, b.ctypes.data, 0))
        return w, b

    def set_weights(self, layer: int, w: np.ndarray, b: np.ndarray) -> None:
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        check("mlc_mlp_weights", load().mlc_mlp_weights(self._h, layer, w.ctypes.data, b.ctypes.data, 1))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        y = np.zeros((x.shape[0], self.nout), np.float32)
        check("mlc_mlp_forward", load().mlc_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def grad(self, x: np.ndarray, t: np.ndarray) -> float:
        x = np.ascontiguousarray(x, np.float32)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        check("mlc_mlp_grad", load().mlc_mlp_grad(self._h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                                  C.byref(loss)))
        return loss.value

    def gradients(self, layer: int):
        i, o = self.dims(layer)
        gw = np.zeros((i, o), np.float32)
        gb = np.zeros(o, np.float32)
        check("mlc_mlp_gradients", load().mlc_mlp_gradients(self._h, layer, gw.ctypes.data, gb.ctypes.data))
        return gw, gb

    def adam(self, lr: float) -> None:
        check("mlc_mlp_adam", load().mlc_mlp_adam(self._h, lr))

    def close(self) -> None:
        if self._h is not None:
            load().mlc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gemm_bf16(a: np.ndarray, b: np.ndarray, a_t: bool = False, b_t: bool = False) -> np.ndarray:
    """C = op(a) . op(b) on the MFMA GEMM kernel (operands rounded to BF16)."""
    a = np.ascontiguousarray(a, np.float32)
    b = np.ascontiguousarray(b, np.float32)
    if a_t:
        K, M = a.shape
        sam, sak = 1, M
    else:
        M, K = a.shape
        sam, sak = K, 1
    if b_t:
        N, K2 = b.shape
        sbk, sbn = 1, K2
    else:
        K2, N = b.shape
        sbk, sbn = N, 1
    assert K == K2
    c = np.zeros((M, N), np.float32)
    check("mlc_gemm_bf16_host", load().mlc_gemm_bf16_host(a.ctypes.data, sam, sak, a.size, b.ctypes.data, sbk, sbn,
                                                          b.size, M, N, K, c.ctypes.data))
    return c


def to_bf16(x: np.ndarray) -> np.ndarray:
    """Round float32 values to bfloat16 (nearest even), returned as float32."""
    x = np.ascontiguousarray(x, np.float32)
    u = x.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)
