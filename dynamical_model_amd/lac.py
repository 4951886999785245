"""Learned acoustic corrector: the column-stencil MLP engine (C ABI mlc_mlp_*).

The network maps a 3 x 3 stencil of columns -- u, v, w, Exner, theta, qv at
kz levels each -- to corrections of u, v, w and Exner at kz levels, trained
on pairs generated by the plain GPU core.  The reference has no learned
component; this is new.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import MLC_LAC_OFF, MLC_LAC_ON, MLC_LAC_REDUCED, LacConfig, LacStats, check, load

MODES = {"off": MLC_LAC_OFF, "on": MLC_LAC_ON, "reduced": MLC_LAC_REDUCED}


def nin_for(kz: int) -> int:
    """Inputs of the corrector: 9 stencil columns x roundup(6 kz, 16)."""
    return 9 * ((6 * kz + 15) // 16 * 16)


def nout_for(kz: int) -> int:
    return 4 * kz


def feature_index(kz: int):
    """(fl, p) of every input: fl = field*kz + level-1 (fields u, v, w at
    levels 2..kz+1, Exner, theta, qv; fl >= 6 kz is zero padding) and the
    stencil position p, (dj, di) = (p%3 - 1, p//3 - 1)."""
    o = np.arange(nin_for(kz))
    c = o >> 3
    cc = c % 18
    return (c // 18) * 16 + (cc // 9) * 8 + (o & 7), cc % 9


class MLP:
    """nin -> nhid -> nhid -> nout with ReLU, BF16 MFMA, FP32 master weights."""

    def __init__(self, nin: int, nhid: int, nout: int, seed: int = 1, device: int = 0, _borrowed=None):
        self.nin, self.nhid, self.nout = nin, nhid, nout
        if _borrowed is not None:  # a network owned by a model handle
            self._h = C.c_void_p(_borrowed)
            self._owned = False
            return
        h = C.c_void_p()
        check("mlc_mlp_create", load().mlc_mlp_create(nin, nhid, nout, seed, device, C.byref(h)))
        self._h = h
        self._owned = True

    @property
    def handle(self):
        return self._h

    def dims(self, layer: int) -> tuple[int, int]:
        return ([self.nin, self.nhid, self.nhid][layer], [self.nhid, self.nhid, self.nout][layer])

    def _layer(self, layer: int) -> tuple[int, int]:
        if layer not in (0, 1, 2):
            raise ValueError(f"layer {layer}: the network has layers 0, 1, 2")
        return self.dims(layer)

    def _rows(self, x: np.ndarray, what: str) -> np.ndarray:
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 2 or x.shape[1] != self.nin:
            raise ValueError(f"{what}: inputs of shape {x.shape}, expected (rows, {self.nin})")
        return x

    def get_weights(self, layer: int):
        i, o = self._layer(layer)
        w = np.zeros((i, o), np.float32)
        b = np.zeros(o, np.float32)
        check("mlc_mlp_weights", load().mlc_mlp_weights(self._h, layer, w.ctypes.data, b.ctypes.data, 0))
        return w, b

    def set_weights(self, layer: int, w: np.ndarray, b: np.ndarray) -> None:
        i, o = self._layer(layer)
        w = np.ascontiguousarray(w, np.float32)
        b = np.ascontiguousarray(b, np.float32)
        if w.shape != (i, o) or b.shape != (o,):
            raise ValueError(f"layer {layer}: weights {w.shape} and biases {b.shape}, expected ({i}, {o}) and ({o},)")
        check("mlc_mlp_weights", load().mlc_mlp_weights(self._h, layer, w.ctypes.data, b.ctypes.data, 1))

    def forward(self, x: np.ndarray) -> np.ndarray:
        x = self._rows(x, "forward")
        y = np.zeros((x.shape[0], self.nout), np.float32)
        check("mlc_mlp_forward", load().mlc_mlp_forward(self._h, x.ctypes.data, x.shape[0], y.ctypes.data))
        return y

    def forward_noalloc(self, x_dev: int, rows: int, y_dev: int) -> None:
        """Y = network(X) for device arrays (addresses of fp32 [rows x nin] and
        [rows x nout]) without any device allocation by the library."""
        check("mlc_mlp_forward_noalloc", load().mlc_mlp_forward_noalloc(self._h, C.c_void_p(x_dev), int(rows),
                                                                        C.c_void_p(y_dev)))

    def grad(self, x: np.ndarray, t: np.ndarray) -> float:
        x = self._rows(x, "grad")
        t = np.ascontiguousarray(t, np.float32)
        if t.shape != (x.shape[0], self.nout):
            raise ValueError(f"grad: targets of shape {t.shape}, expected ({x.shape[0]}, {self.nout})")
        loss = C.c_double()
        check("mlc_mlp_grad", load().mlc_mlp_grad(self._h, x.ctypes.data, t.ctypes.data, x.shape[0],
                                                  C.byref(loss)))
        return loss.value

    def gradients(self, layer: int):
        i, o = self._layer(layer)
        gw = np.zeros((i, o), np.float32)
        gb = np.zeros(o, np.float32)
        check("mlc_mlp_gradients", load().mlc_mlp_gradients(self._h, layer, gw.ctypes.data, gb.ctypes.data))
        return gw, gb

    def adam(self, lr: float) -> None:
        check("mlc_mlp_adam", load().mlc_mlp_adam(self._h, lr))

    def close(self) -> None:
        if self._h is not None:
            if self._owned:
                load().mlc_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def default_config(**overrides) -> LacConfig:
    c = LacConfig()
    check("mlc_lac_default_config", load().mlc_lac_default_config(C.byref(c)))
    for k, v in overrides.items():
        setattr(c, k, v)
    return c


class Corrector:
    """The learned acoustic corrector attached to a Moloch handle."""

    def __init__(self, model, config: LacConfig | None = None, **overrides):
        self.model = model
        cfg = config if config is not None else default_config(**overrides)
        check("mlc_lac_create", load().mlc_lac_create(model._h, C.byref(cfg)))
        self.config = cfg

    def collect(self, nsteps: int) -> None:
        """nsteps ordinary steps that also record training pairs."""
        check("mlc_lac_collect", load().mlc_lac_collect(self.model._h, int(nsteps)))

    def train(self, iters: int, batch: int = 512, lr: float = 1e-3) -> float:
        loss = C.c_double()
        check("mlc_lac_train", load().mlc_lac_train(self.model._h, int(iters), int(batch), float(lr),
                                                    C.byref(loss)))
        return loss.value

    def mode(self, mode: str) -> None:
        check("mlc_lac_mode", load().mlc_lac_mode(self.model._h, MODES[mode]))

    def stats(self) -> dict:
        s = LacStats()
        check("mlc_lac_get_stats", load().mlc_lac_get_stats(self.model._h, C.byref(s)))
        return {"samples": s.samples, "learned_phases": s.learned_phases, "plain_phases": s.plain_phases,
                "checks": s.checks, "fallback": bool(s.fallback), "last_loss": s.last_loss,
                "ratio": list(s.ratio), "rel": list(s.rel), "worst_ratio": s.worst_ratio,
                "worst_rel": s.worst_rel, "retrains": s.retrains, "collect_phases": s.collect_phases,
                "failed_checks": s.failed_checks}

    def dataset(self, first: int, n: int):
        """Rows of the training set: (X raw inputs, T targets, sampled columns);
        the widths follow the handle's level count."""
        kz = self.model.kz
        if first < 0 or n < 0:
            raise ValueError("dataset: negative row range")
        X = np.zeros((n, nin_for(kz)), np.float32)
        T = np.zeros((n, nout_for(kz)), np.float32)
        cols = np.zeros(n, np.int32)
        check("mlc_lac_dataset", load().mlc_lac_dataset(self.model._h, int(first), int(n), X.ctypes.data,
                                                        T.ctypes.data, cols.ctypes.data))
        return X, T, cols

    def network(self) -> MLP:
        """The handle's network (dimensions from the handle's level count)."""
        ptr = load().mlc_lac_network(self.model._h)
        if not ptr:
            raise RuntimeError("the handle has no corrector network")
        kz = self.model.kz
        return MLP(nin_for(kz), self.config.nhid, nout_for(kz), _borrowed=ptr)


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
