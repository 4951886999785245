"""ctypes binding of the C ABI declared in include/moloch_hip.h.

The shared library is built in-tree by ``dynamical_model_amd.build``; there
is no fallback: if it is missing, loading raises.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libmoloch_hip.so")

# error codes
MLC_OK, MLC_EINVAL, MLC_EHIP, MLC_ESTATE, MLC_ENOMEM = 0, 1, 2, 3, 4
# precision
MLC_FP64, MLC_FP32 = 0, 1
# case kinds
MLC_CASE_STANDARD, MLC_CASE_REFERENCE = 0, 1

FIELDS = [
    "u", "v", "w", "pai", "t", "tetav", "qv", "qc", "qi", "qr", "qs",
    "tvirt", "p", "rho", "qsat", "ux", "vx", "ps",
    "fmz", "fmzf", "zeta", "hx", "hy", "msfu", "msfv", "msfx", "ulat", "vlat",
]
FIELD_ID = {name: i for i, name in enumerate(FIELDS)}
FIELDS_2D = {"ps", "hx", "hy", "msfu", "msfv", "msfx", "ulat", "vlat"}
FIELDS_KP1 = {"w", "fmzf"}
STATIC_FIELDS = ["fmz", "fmzf", "zeta", "hx", "hy", "msfu", "msfv", "msfx", "ulat", "vlat"]
STATE_FIELDS = [f for f in FIELDS if f not in STATIC_FIELDS]
LEVELS = ["zita", "zitah", "ffilt"]


def nlev(name: str, kz: int) -> int:
    if name in FIELDS_2D:
        return 1
    return kz + 1 if name in FIELDS_KP1 else kz


class Dims(C.Structure):
    _fields_ = [("jx", C.c_int), ("iy", C.c_int), ("kz", C.c_int), ("nqx", C.c_int)]


class Params(C.Structure):
    _fields_ = [
        ("dt", C.c_double), ("dx", C.c_double), ("dzita", C.c_double), ("ztop", C.c_double),
        ("zh", C.c_double), ("a0", C.c_double), ("anu2", C.c_double),
        ("nadv", C.c_int), ("nsound", C.c_int), ("iqfrst", C.c_int),
        ("rotllr", C.c_int), ("divfilter", C.c_int), ("halo_mpi", C.c_int),
        ("fulleq", C.c_int), ("divdamp", C.c_int), ("vadvtwice", C.c_int),
        ("filterpai", C.c_int), ("filtertheta", C.c_int),
        ("precision", C.c_int), ("device", C.c_int), ("unfused", C.c_int),
    ]


class Diag(C.Structure):
    _fields_ = [
        ("ps_max", C.c_double), ("ps_min", C.c_double), ("u_max", C.c_double),
        ("u_min", C.c_double), ("v_max", C.c_double), ("v_min", C.c_double),
        ("cfl_violation", C.c_int),
    ]


class Case(C.Structure):
    _fields_ = [
        ("kind", C.c_int), ("jx", C.c_int), ("iy", C.c_int), ("kz", C.c_int),
        ("ztop", C.c_double), ("zh", C.c_double), ("a0", C.c_double),
        ("nhills", C.c_int),
        ("hill_h", C.c_double), ("hill_hw", C.c_double),
        ("hill_hmin", C.c_double), ("hill_hmax", C.c_double),
        ("hill_hwmin", C.c_double), ("hill_hwmax", C.c_double),
        ("tsfc", C.c_double), ("lapse", C.c_double), ("ttrop", C.c_double),
        ("psfc", C.c_double), ("rh", C.c_double),
        ("u0", C.c_double), ("v0", C.c_double), ("jet_amp", C.c_double),
        ("jet_z", C.c_double), ("jet_width", C.c_double),
        ("noise_amp", C.c_double), ("noise_levels", C.c_int),
        ("bubble_amp", C.c_double), ("bubble_rcells", C.c_double),
        ("bubble_z", C.c_double), ("bubble_rz", C.c_double),
        ("zero_coriolis", C.c_int), ("seed", C.c_ulonglong),
    ]


class LacConfig(C.Structure):
    _fields_ = [
        ("nhid", C.c_int), ("samples_per_phase", C.c_int), ("capacity", C.c_int),
        ("chunk_rows", C.c_int), ("check_every", C.c_int), ("check_tol", C.c_double),
        ("seed", C.c_ulonglong), ("engine", C.c_int),
        ("rel_tol", C.c_double), ("retrain_steps", C.c_int), ("retrain_iters", C.c_int),
        ("retrain_batch", C.c_int), ("retrain_lr", C.c_double), ("max_retrains", C.c_int),
    ]


class LacStats(C.Structure):
    _fields_ = [
        ("samples", C.c_longlong), ("learned_phases", C.c_longlong),
        ("plain_phases", C.c_longlong), ("checks", C.c_longlong), ("fallback", C.c_int),
        ("last_loss", C.c_double), ("ratio", C.c_double * 4), ("rel", C.c_double * 4),
        ("worst_ratio", C.c_double), ("worst_rel", C.c_double), ("retrains", C.c_longlong),
        ("collect_phases", C.c_longlong), ("failed_checks", C.c_longlong),
    ]


MLC_LAC_OFF, MLC_LAC_ON, MLC_LAC_REDUCED = 0, 1, 2

EXPORTS = {
    "mlc_default_params": (C.c_int, [C.POINTER(Params)]),
    "mlc_create": (C.c_int, [C.POINTER(Dims), C.POINTER(Params), C.POINTER(C.c_void_p)]),
    "mlc_set_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mlc_get_field": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
    "mlc_set_levels": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_int]),
    "mlc_init": (C.c_int, [C.c_void_p]),
    "mlc_step": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.POINTER(Diag)]),
    "mlc_sync": (C.c_int, [C.c_void_p]),
    "mlc_debug_poison_halos": (C.c_int, [C.c_void_p, C.c_int]),
    "mlc_destroy": (C.c_int, [C.c_void_p]),
    "mlc_last_error": (C.c_char_p, []),
    "mlc_case_config": (C.c_int, [C.c_int, C.POINTER(Case)]),
    "mlc_case_params": (C.c_int, [C.POINTER(Case), C.POINTER(Params)]),
    "mlc_case_build": (C.c_int, [C.POINTER(Case), C.POINTER(C.c_void_p), C.POINTER(C.c_void_p)]),
    "mlc_case_load": (C.c_int, [C.c_void_p, C.POINTER(Case)]),
    "mlc_case_load_device": (C.c_int, [C.c_void_p, C.POINTER(Case)]),
    "mlc_mlp_create": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_ulonglong, C.c_int, C.POINTER(C.c_void_p)]),
    "mlc_mlp_destroy": (C.c_int, [C.c_void_p]),
    "mlc_mlp_weights": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p, C.c_int]),
    "mlc_mlp_gradients": (C.c_int, [C.c_void_p, C.c_int, C.c_void_p, C.c_void_p]),
    "mlc_mlp_forward": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "mlc_mlp_forward_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]),
    "mlc_mlp_grad": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_double)]),
    "mlc_mlp_adam": (C.c_int, [C.c_void_p, C.c_double]),
    "mlc_gemm_bf16_host": (C.c_int, [C.c_void_p, C.c_long, C.c_long, C.c_long, C.c_void_p, C.c_long, C.c_long,
                                     C.c_long, C.c_int, C.c_int, C.c_int, C.c_void_p]),
    "mlc_lac_default_config": (C.c_int, [C.POINTER(LacConfig)]),
    "mlc_lac_create": (C.c_int, [C.c_void_p, C.POINTER(LacConfig)]),
    "mlc_lac_collect": (C.c_int, [C.c_void_p, C.c_int]),
    "mlc_lac_train": (C.c_int, [C.c_void_p, C.c_int, C.c_int, C.c_double, C.POINTER(C.c_double)]),
    "mlc_lac_mode": (C.c_int, [C.c_void_p, C.c_int]),
    "mlc_lac_get_stats": (C.c_int, [C.c_void_p, C.POINTER(LacStats)]),
    "mlc_lac_network": (C.c_void_p, [C.c_void_p]),
    "mlc_lac_dataset": (C.c_int, [C.c_void_p, C.c_long, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p]),
    "mlc_nc_create": (C.c_int, [C.c_char_p, C.c_char_p, C.c_int, C.c_int, C.c_int, C.POINTER(C.c_void_p)]),
    "mlc_nc_append": (C.c_int, [C.c_void_p, C.c_void_p]),
    "mlc_nc_records": (C.c_int, [C.c_void_p]),
    "mlc_nc_close": (C.c_int, [C.c_void_p]),
    "mlc_output_open": (C.c_int, [C.c_void_p, C.c_char_p, C.POINTER(C.c_void_p)]),
    "mlc_output_write": (C.c_int, [C.c_void_p]),
    "mlc_output_flush": (C.c_int, [C.c_void_p]),
    "mlc_output_close": (C.c_int, [C.c_void_p]),
    "mlc_profile": (C.c_int, [C.c_void_p, C.c_int]),
    "mlc_profile_kinds": (C.c_int, []),
    "mlc_profile_get": (C.c_int, [C.c_void_p, C.c_int, C.POINTER(C.c_char_p), C.POINTER(C.c_double),
                                  C.POINTER(C.c_longlong), C.POINTER(C.c_double)]),
}

_lib = None


def load() -> C.CDLL:
    """Load the HIP library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m dynamical_model_amd.build`"
            )
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in EXPORTS.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class MolochError(RuntimeError):
    """A non-zero status returned by the C ABI."""

    def __init__(self, func: str, code: int, msg: str):
        super().__init__(f"{func} failed with code {code}: {msg}")
        self.code = code


def check(func: str, rc: int) -> None:
    if rc != MLC_OK:
        raise MolochError(func, rc, load().mlc_last_error().decode(errors="replace"))
