"""Python host mirror of the reference's mod_moloch interface.

The reference drives the timestep through the Fortran module mod_moloch
(nompi/mod_moloch.F90:93): ``allocate_moloch`` (:183-217), ``init_moloch``
(:219-268) and ``moloch(istep, lprint)`` (:272-421), with the state in module
globals.  :class:`Moloch` keeps those three names and meanings and holds the
device state behind the C ABI of include/moloch_hip.h.

Host arrays use the canonical layout of this package: a numpy array of shape
``(nlev, iy+4, jx+4)`` (C order), i.e. the Fortran array
``(-1:jx+2, -1:iy+2, 1:nlev)`` with j fastest; element ``(j, i, k)`` is
``a[k-1, i+1, j+1]``.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import FIELD_ID, FIELDS, LEVELS, MLC_FP32, MLC_FP64, check, load, nlev


def canonical_shape(name: str, jx: int, iy: int, kz: int) -> tuple[int, ...]:
    n = nlev(name, kz)
    if n == 1:
        return (iy + 4, jx + 4)
    return (n, iy + 4, jx + 4)


def interior(a: np.ndarray) -> np.ndarray:
    """View of the interior points (j=1..jx, i=1..iy) of a canonical array."""
    return a[..., 2:-2, 2:-2]


def default_params(**overrides) -> _lib.Params:
    p = _lib.Params()
    check("mlc_default_params", load().mlc_default_params(C.byref(p)))
    for k, v in overrides.items():
        if not hasattr(p, k):
            raise KeyError(k)
        setattr(p, k, v)
    return p


def case_config(config: int, **overrides) -> _lib.Case:
    """Defaults of benchmark configuration ``config`` (0..4)."""
    c = _lib.Case()
    check("mlc_case_config", load().mlc_case_config(int(config), C.byref(c)))
    for k, v in overrides.items():
        if not hasattr(c, k):
            raise KeyError(k)
        setattr(c, k, v)
    return c


def case_params(case: _lib.Case, **overrides) -> _lib.Params:
    p = default_params()
    check("mlc_case_params", load().mlc_case_params(C.byref(case), C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


@dataclass
class HostState:
    """Fields (canonical layout) and level arrays of one run."""

    jx: int
    iy: int
    kz: int
    fields: dict
    levels: dict

    def copy(self) -> "HostState":
        return HostState(self.jx, self.iy, self.kz,
                         {k: v.copy() for k, v in self.fields.items()},
                         {k: v.copy() for k, v in self.levels.items()})


def build_case(case: _lib.Case) -> HostState:
    """Geometry and initial state of ``case``, built by the library on the host."""
    jx, iy, kz = case.jx, case.iy, case.kz
    fields = {n: np.zeros(canonical_shape(n, jx, iy, kz)) for n in FIELDS}
    levels = {"zita": np.zeros(kz + 1), "zitah": np.zeros(kz), "ffilt": np.zeros(kz)}
    fp = (C.c_void_p * len(FIELDS))(*[fields[n].ctypes.data for n in FIELDS])
    lp = (C.c_void_p * 3)(*[levels[n].ctypes.data for n in LEVELS])
    check("mlc_case_build", load().mlc_case_build(C.byref(case), fp, lp))
    return HostState(jx, iy, kz, fields, levels)


class Moloch:
    """One MOLOCH run on a GPU: the module mod_moloch of the reference."""

    def __init__(self):
        self._h = None

    # ---------------------------------------------------------- reference API
    def allocate_moloch(self, jx: int, iy: int, kz: int, params: _lib.Params) -> None:
        """allocate_moloch: create the device state (nompi/mod_moloch.F90:183-217)."""
        self.close()
        self.jx, self.iy, self.kz = jx, iy, kz
        self.params = params
        dims = _lib.Dims(jx, iy, kz, 5)
        h = C.c_void_p()
        check("mlc_create", load().mlc_create(C.byref(dims), C.byref(params), C.byref(h)))
        self._h = h

    def init_moloch(self) -> None:
        """init_moloch: derived coefficients on the device (:219-268)."""
        check("mlc_init", load().mlc_init(self._h))

    def moloch(self, istep: int, lprint: bool = False) -> _lib.Diag | None:
        """moloch(istep, lprint): one step (:272-421); returns the diagnostics if lprint."""
        d = _lib.Diag()
        check("mlc_step", load().mlc_step(self._h, int(istep), 1 if lprint else 0, C.byref(d)))
        return d if lprint else None

    # --------------------------------------------------------------- transfers
    def set_field(self, name: str, a: np.ndarray, lb=None) -> None:
        a = np.ascontiguousarray(a, dtype=np.float64)
        lb, ub = self._bounds(name, a.shape, lb)
        check("mlc_set_field", load().mlc_set_field(self._h, FIELD_ID[name], a.ctypes.data, lb, ub))

    def get_field(self, name: str, out: np.ndarray | None = None, lb=None) -> np.ndarray:
        if out is None:
            out = np.zeros(canonical_shape(name, self.jx, self.iy, self.kz))
        lb, ub = self._bounds(name, out.shape, lb)
        check("mlc_get_field", load().mlc_get_field(self._h, FIELD_ID[name], out.ctypes.data, lb, ub))
        return out

    def set_levels(self, name: str, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, dtype=np.float64)
        check("mlc_set_levels", load().mlc_set_levels(self._h, LEVELS.index(name), v.ctypes.data, v.size))

    def _bounds(self, name, shape, lb):
        rank = 2 if len(shape) == 2 else 3
        if lb is None:
            lb = (-1, -1, 1)[:rank]
        # numpy shape is (nlev, ni, nj) -> Fortran extents (nj, ni, nlev)
        ext = tuple(reversed(shape))
        ub = tuple(l + e - 1 for l, e in zip(lb, ext))
        return (C.c_int * 3)(*lb, *([1] * (3 - rank))), (C.c_int * 3)(*ub, *([1] * (3 - rank)))

    def load_state(self, hs: HostState) -> None:
        for name in FIELDS:
            self.set_field(name, hs.fields[name])
        for name in LEVELS:
            self.set_levels(name, hs.levels[name])

    # ------------------------------------------------------ checkpoint/resume
    # What a step reads of the state it continues from: the prognostic fields
    # and the diagnostics it recomputes, as whole canonical arrays (halos
    # included, so the nompi tree's stale halo seam carries over too).  The
    # geometry is not part of it: a restored handle is set up from the same
    # case first.  Values pass as FP64, which holds the FP32 state exactly.
    STATE_FIELDS = ["u", "v", "w", "pai", "t", "tetav", "qv", "qc", "qi", "qr", "qs",
                    "tvirt", "p", "rho", "qsat", "ux", "vx", "ps"]

    def save_state(self, path: str, step: int) -> None:
        """Raw dump of the state after step ``step`` (an .npz file)."""
        self.sync()
        arrays = {f: self.get_field(f) for f in self.STATE_FIELDS}
        np.savez(path, step=np.int64(step), jx=self.jx, iy=self.iy, kz=self.kz,
                 precision=np.int64(self.params.precision), **arrays)

    def restore_state(self, path: str) -> int:
        """Load a dump written by save_state into this handle (same grid and
        geometry, after init_moloch or a case load); returns its step."""
        with np.load(path) as d:
            if (int(d["jx"]), int(d["iy"]), int(d["kz"])) != (self.jx, self.iy, self.kz):
                raise ValueError(f"{path}: grid {int(d['jx'])}x{int(d['iy'])}x{int(d['kz'])} differs from the "
                                 f"handle's {self.jx}x{self.iy}x{self.kz}")
            for f in self.STATE_FIELDS:
                self.set_field(f, d[f])
            return int(d["step"])

    def load_case(self, case: _lib.Case) -> None:
        check("mlc_case_load", load().mlc_case_load(self._h, C.byref(case)))

    def load_case_device(self, case: _lib.Case) -> None:
        """The case built on the device (3-D fields) instead of the host."""
        check("mlc_case_load_device", load().mlc_case_load_device(self._h, C.byref(case)))

    def sync(self) -> None:
        check("mlc_sync", load().mlc_sync(self._h))

    def poison_halos(self, enable: bool = True) -> None:
        """Debug: NaN into every state and scratch halo before each step."""
        check("mlc_debug_poison_halos", load().mlc_debug_poison_halos(self._h, 1 if enable else 0))

    def profile(self, enable: bool) -> None:
        """Bracket every kernel launch with HIP events (and clear the totals)."""
        check("mlc_profile", load().mlc_profile(self._h, 1 if enable else 0))

    def profile_report(self) -> dict:
        """{kernel kind: (total ms, launches, algorithmic bytes)} since profile(True)."""
        out = {}
        for k in range(load().mlc_profile_kinds()):
            name = C.c_char_p()
            ms = C.c_double()
            n = C.c_longlong()
            b = C.c_double()
            check("mlc_profile_get", load().mlc_profile_get(self._h, k, C.byref(name), C.byref(ms),
                                                            C.byref(n), C.byref(b)))
            if n.value:
                out[name.value.decode()] = (ms.value, n.value, b.value)
        return out

    def close(self) -> None:
        if self._h is not None:
            load().mlc_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def run_case(hs: HostState, params: _lib.Params, nsteps: int, fields=None) -> dict:
    """Upload ``hs``, take ``nsteps`` steps on the GPU and return the fields."""
    m = Moloch()
    try:
        m.allocate_moloch(hs.jx, hs.iy, hs.kz, params)
        m.load_state(hs)
        m.init_moloch()
        for n in range(1, nsteps + 1):
            m.moloch(n, False)
        m.sync()
        names = fields or FIELDS
        return {n: m.get_field(n) for n in names}
    finally:
        m.close()


__all__ = [
    "Moloch", "HostState", "build_case", "case_config", "case_params", "default_params",
    "canonical_shape", "interior", "run_case", "MLC_FP64", "MLC_FP32",
]
