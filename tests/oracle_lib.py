"""ctypes wrapper of the CPU oracle (oracle/libmoloch_oracle.so).

Test infrastructure: the oracle is the checker the HIP implementation is
compared against.  Arrays use the package's canonical layout
(nlev, iy+4, jx+4), which is the oracle's own layout.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "libmoloch_oracle.so")
# the same source built with loop-level OpenMP threads (the all-core CPU
# baseline); bit-identical to the serial build
ORACLE_LIB_OMP = os.path.join(ORACLE_DIR, "libmoloch_oracle_omp.so")

_D = C.POINTER(C.c_double)

_PTRS_1D = ["zita", "zitah", "ffilt", "gzitak", "gzitakh", "xknu"]
_PTRS_2D = ["hx", "hy", "mu", "mv", "mx", "ulat", "vlat", "ps", "rmu", "rmv", "mx2", "coru", "corv"]
_PTRS_3D = ["fmz", "fmzf", "zeta", "u", "v", "w", "pai", "t", "tetav", "tvirt", "p", "rho", "qsat",
            "ux", "vx", "qx", "s", "deltaw", "zdiv2", "wx", "wz", "p0", "ud", "vd", "qwltot",
            "qwitot", "pf", "tf", "p2d", "wwkw", "wfw", "zpby", "zpbw"]
_KP1 = {"fmzf", "w", "s", "deltaw", "wwkw", "wfw"}


class OraState(C.Structure):
    _fields_ = (
        [("nj", C.c_int), ("ni", C.c_int), ("nk", C.c_int)]
        + [(n, C.c_double) for n in ("dt", "dx", "dzita", "ztop", "zh", "a0", "anu2")]
        + [(n, C.c_int) for n in ("nadv", "nsound", "nqx", "iqfrst", "rotllr", "divfilter",
                                  "halo_mpi", "do_fulleq", "do_divdamp", "do_vadvtwice",
                                  "do_filterpai", "do_filtertheta")]
        + [(n, _D) for n in _PTRS_1D]
        + [(n, _D) for n in ["hx", "hy", "mu", "mv", "mx", "ulat", "vlat", "ps",
                              "rmu", "rmv", "mx2", "coru", "corv"]]
        + [(n, _D) for n in ["fmz", "fmzf", "zeta", "u", "v", "w", "pai", "t", "tetav", "tvirt",
                              "p", "rho", "qsat", "ux", "vx", "qx", "s", "deltaw", "zdiv2",
                              "wx", "wz", "p0", "ud", "vd", "qwltot", "qwitot", "pf", "tf",
                              "p2d", "wwkw", "wfw", "zpby", "zpbw"]]
        + [("diag", C.c_double * 6), ("cfl_violation", C.c_int)]
    )


class OraConstants(C.Structure):
    _fields_ = [(n, C.c_double) for n in (
        "egrav", "rgas", "rwat", "cpd", "cvd", "rdrcv", "cpovr", "rovcp", "p00", "ep1", "ep2",
        "wlhv", "tzero", "lrate", "eomeg2", "degrad", "mathpi", "regrav", "earthrad", "raddeg")]


_libs = {}


def build_oracle() -> None:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)


def lib(threaded: bool = False) -> C.CDLL:
    if threaded not in _libs:
        path = ORACLE_LIB_OMP if threaded else ORACLE_LIB
        # a scratch build of the oracle with a planted mistake (the mutation
        # test of the pins runs the pins against it in a child process)
        path = os.environ.get("MLC_ORACLE_MUTANT", path)
        if not os.path.exists(path):
            build_oracle()
        L = C.CDLL(path)
        P = C.POINTER(OraState)
        for name, args in {
            "ora_init_moloch": [P], "ora_reset_tendencies": [P], "ora_prologue": [P],
            "ora_sound": [P, C.c_double], "ora_divdamp": [P, C.c_double],
            "ora_advection": [P, C.c_double], "ora_wafone": [P, _D, C.c_double],
            "ora_uvstagtox": [P], "ora_wstagtox": [P], "ora_xtouvstag": [P], "ora_xtowstag": [P],
            "ora_epilogue": [P], "ora_diagnostics": [P], "ora_filtpai": [P], "ora_filttheta": [P],
            "ora_divergence_filter": [P], "ora_moloch": [P, C.c_int, C.c_int],
            "ora_thomas_column": [C.c_int, _D, _D, _D, C.c_double, C.c_double, _D],
            "ora_get_constants": [C.POINTER(OraConstants)],
            "ora_substep": [P, C.c_int, C.c_double], "ora_sound_eq10": [P],
            "ora_sound_hdiv": [P, C.c_double], "ora_sound_vdiv": [P, C.c_double],
            "ora_sound_wsolve": [P, C.c_int, C.c_double], "ora_sound_exner": [P, C.c_double],
            "ora_sound_momentum": [P, C.c_double],
            "ora_w_coefficients": [P, C.c_int, C.c_int, C.c_int, C.c_double, _D, _D, _D, _D],
            "ora_fulleq_point": [C.c_double, C.c_double, C.c_double, C.c_double, _D, _D],
        }.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = None
        for name in ("ora_pfesat", "ora_pfwsat"):
            f = getattr(L, name)
            f.argtypes = [C.c_double, C.c_double]
            f.restype = C.c_double
        L.ora_gzita.argtypes = [C.c_double, C.c_double, C.c_double]
        L.ora_gzita.restype = C.c_double
        _libs[threaded] = L
    return _libs[threaded]


def constants() -> dict:
    c = OraConstants()
    lib().ora_get_constants(C.byref(c))
    return {n: getattr(c, n) for n, _ in OraConstants._fields_}


def _ptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_D)


class Oracle:
    """One oracle run: all module state and scratch as numpy arrays."""

    STATE_MAP = {  # host field name -> oracle array name
        "u": "u", "v": "v", "w": "w", "pai": "pai", "t": "t", "tetav": "tetav", "tvirt": "tvirt",
        "p": "p", "rho": "rho", "qsat": "qsat", "ux": "ux", "vx": "vx", "ps": "ps",
        "fmz": "fmz", "fmzf": "fmzf", "zeta": "zeta", "hx": "hx", "hy": "hy",
        "msfu": "mu", "msfv": "mv", "msfx": "mx", "ulat": "ulat", "vlat": "vlat",
    }
    QX = ["qv", "qc", "qi", "qr", "qs"]

    def __init__(self, jx, iy, kz, params, levels=None, threaded=False):
        self.jx, self.iy, self.kz = jx, iy, kz
        self.L = lib(threaded)
        self.a = {}
        plane = (iy + 4, jx + 4)
        for n in _PTRS_1D:
            self.a[n] = np.zeros(kz + 1)
        for n in _PTRS_2D:
            self.a[n] = np.zeros(plane)
        for n in _PTRS_3D:
            if n == "qx":
                self.a[n] = np.zeros((5, kz) + plane)
            elif n == "p2d":
                self.a[n] = np.zeros(plane)
            else:
                self.a[n] = np.zeros(((kz + 1) if n in _KP1 else kz,) + plane)
        S = OraState()
        S.nj, S.ni, S.nk = jx, iy, kz
        for n in ("dt", "dx", "dzita", "ztop", "zh", "a0", "anu2", "nadv", "nsound", "iqfrst",
                  "rotllr", "divfilter", "halo_mpi"):
            setattr(S, n, getattr(params, n))
        S.nqx = 5
        S.do_fulleq = params.fulleq
        S.do_divdamp = params.divdamp
        S.do_vadvtwice = params.vadvtwice
        S.do_filterpai = params.filterpai
        S.do_filtertheta = params.filtertheta
        for n, arr in self.a.items():
            setattr(S, n, _ptr(arr))
        self.S = S
        if levels is not None:
            self.a["zita"][: kz + 1] = levels["zita"]
            self.a["zitah"][:kz] = levels["zitah"]
            self.a["ffilt"][:kz] = levels["ffilt"]

    @property
    def qx(self):
        return {n: self.a["qx"][i] for i, n in enumerate(self.QX)}

    def load_fields(self, fields: dict) -> None:
        for name, arr in fields.items():
            if name in self.QX:
                self.a["qx"][self.QX.index(name)][...] = arr
            elif name in self.STATE_MAP:
                self.a[self.STATE_MAP[name]][...] = arr

    def field(self, name: str) -> np.ndarray:
        if name in self.QX:
            return self.a["qx"][self.QX.index(name)]
        return self.a[self.STATE_MAP[name]]

    def init_moloch(self):
        self.L.ora_init_moloch(C.byref(self.S))

    def moloch(self, istep=1, lprint=False):
        self.L.ora_moloch(C.byref(self.S), int(istep), 1 if lprint else 0)

    def call(self, name, *args):
        return getattr(self.L, name)(C.byref(self.S), *args)

    def wafone(self, arr: np.ndarray, dta: float):
        self.L.ora_wafone(C.byref(self.S), _ptr(arr), dta)


def oracle_from_host(hs, params, threaded: bool = False) -> Oracle:
    o = Oracle(hs.jx, hs.iy, hs.kz, params, hs.levels, threaded)
    o.load_fields(hs.fields)
    o.init_moloch()
    return o
