"""The step's output files (C ABI mlc_nc_*, mlc_output_*).

The reference writes ps and t as one netCDF classic file per variable with
grid_nc_create / grid_nc_write / grid_nc_destroy
(nompi/mod_mppparam.F90:121-480) from the driver's output(istep)
(nompi/regcm_dynamical_core.F90:202-215).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import check, load


class NcFile:
    """One double record variable over (KTAU, [KZ,] IY, JX)."""

    def __init__(self, path: str, varname: str, nx: int, ny: int, nz: int = 0):
        self.shape = (nz, ny, nx) if nz else (ny, nx)
        h = C.c_void_p()
        check("mlc_nc_create", load().mlc_nc_create(path.encode(), varname.encode(), nx, ny, nz, C.byref(h)))
        self._h = h

    def append(self, values: np.ndarray) -> None:
        v = np.ascontiguousarray(values, np.float64)
        if v.shape != self.shape:
            raise ValueError(f"record shape {v.shape} != {self.shape}")
        check("mlc_nc_append", load().mlc_nc_append(self._h, v.ctypes.data))

    @property
    def records(self) -> int:
        return load().mlc_nc_records(self._h)

    def close(self) -> None:
        if self._h is not None:
            check("mlc_nc_close", load().mlc_nc_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Output:
    """ps.nc and t.nc of a Moloch handle, written without stopping the steps."""

    def __init__(self, model, directory: str):
        h = C.c_void_p()
        check("mlc_output_open", load().mlc_output_open(model._h, directory.encode(), C.byref(h)))
        self._h = h
        self.model = model

    def write(self) -> None:
        """Queue one record of ps and t as they are after the queued steps."""
        check("mlc_output_write", load().mlc_output_write(self._h))

    def flush(self) -> None:
        check("mlc_output_flush", load().mlc_output_flush(self._h))

    def close(self) -> None:
        if self._h is not None:
            check("mlc_output_close", load().mlc_output_close(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
